// gta_kernels.hip -- CDNA4 (gfx950) kernels of the structure-alignment stage.
//
// Every kernel works on one (query, reference) pair per workgroup; pair p is
// (p / nr, p % nr).  Per-pair state lives in HBM (DevState) between kernels,
// so the host never synchronises inside a chunk.
//
// Reference correspondence (kernel -> reference functions it replaces):
//   k_secstr          MpSecStr::ssk_kernel_helper            MpSecStr.cpp:20-110
//   k_init            MpStage1::Preinitialize1Kernel         MpStage1.cpp:28-170
//   k_findfrag        MpStage1::FindFragKernel               MpStage1.cpp:275-440
//   k_refine_gapless  MpStage1::RefineFragInitKernel         MpStage1.cpp:470-640
//   k_dp              MpDPHub::ExecDPwBtck128xKernel,        MpDPHub.cpp:73-388
//                     ExecDPTFMSSwBtck128xKernel,            MpDPHub.cpp:687-984
//                     BtckToMatched128xKernel,               MpDPHub.cpp:1500-1610
//                     ConstrainedBtckToMatched128xKernel     MpDPHub.cpp:1620-1810
//   k_refine_dp       MpStage1::RefineFragDPKernel (+ the    MpStage1.cpp:780-960,1070-1215
//                     DPRefine bookkeeping kernels),
//                     MpStageFin Final...Phase1/Phase2,      MpStageFin.cpp:44-470
//                     Production...Phase1/Phase2             MpStageFin.cpp:747-1470
//   k_prod_aln        MpDPHub::ProductionMatchToAlignment... MpDPHub.cpp:1919-2142
//   k_setlow          MpStage1::SetLowScoreConvergenceFlag   MpStage1.cpp:1175-1215
//   k_revert          MpStageFin::RevertTfmMatricesKernel    MpStageFin.cpp:1835-1880
//
// Precision P: 0 = FP32 throughout (the reference's arithmetic), 1 = mixed
// (per-lane compensated FP32 accumulation combined in FP64, FP64 3x3 solve,
// FP64 RMSD).  DP cell arithmetic is FP32 in both and rounds exactly like the
// reference's scalar code.

#include <cmath>
#include <vector>

#include "gta_device.h"
#include "gta_gonnet.h"
#include "gta_kabsch.h"
#include "gta_slots.h"

namespace gta {

// ---- d0 / d8 (stagecnsts.cuh:48-175) -----------------------------------------
// cbrtf(L - 15) and powf(L, 0.3) per length L are evaluated once on the host,
// with the C library the reference's CPU path uses: the device's own
// cbrtf/powf differ from it in the last bit for some lengths, and d0 enters
// every DP cell score, so a last-bit difference could change an alignment.
constexpr int MAXLEN = 65535;
__device__ float g_cbrt_lm15[MAXLEN + 1];   // cbrtf((float)L - 15.f)
__device__ float g_pow03[MAXLEN + 1];       // powf((float)L, 0.3f)

__device__ __forceinline__ int GetLnormI(int q, int r) { return min(q, r); }
__device__ __forceinline__ float GetD82(int q, int r) {
    float d8 = 1.5f * g_pow03[GetLnormI(q, r)] + 3.5f;
    return d8 * d8;
}
__device__ __forceinline__ float GetD0(int q, int r) {
    const int L = GetLnormI(q, r);
    float d0 = 0.168f;
    if ((float)L > 19.f) d0 = 1.24f * g_cbrt_lm15[L] - 1.8f;
    return d0 + 0.8f;
}
__device__ __forceinline__ float GetD0finP(int q, int r) {
    const int L = GetLnormI(q, r);
    float d0 = 0.5f;
    if ((float)L > 21.f) d0 = 1.24f * g_cbrt_lm15[L] - 1.8f;
    if (d0 < 0.5f) d0 = 0.5f;
    return d0;
}
__device__ __forceinline__ float GetD0finNA(int q, int r) {
    float l = (float)GetLnormI(q, r), d0 = 0.3f;
    if (30.f <= l) d0 = 0.6f * sqrtf(l - 0.5f) - 2.5f;
    else if (23.f < l) d0 = 0.7f;
    else if (19.f < l) d0 = 0.6f;
    else if (15.f < l) d0 = 0.5f;
    else if (11.f < l) d0 = 0.4f;
    return fmaxf(0.3f, d0);
}
__device__ __forceinline__ float GetD0fin(int q, int r, int na) {
    return na ? GetD0finNA(q, r) : GetD0finP(q, r);
}
__device__ __forceinline__ float GetD0s(float d0) { return fmaxf(fminf(d0, 8.f), 4.5f); }

constexpr int RWAVES = REFINE_WAVES;      // waves per workgroup in the refinement kernels
constexpr int RTHREADS = RWAVES * GTA_WAVE;
constexpr int FTHREADS = SEARCH_WAVES * GTA_WAVE;   // threads per workgroup of the seed search
// The FP32 refinements run 16 waves per CU in workgroups of W waves, held to
// 128 VGPRs, with alignments staged in LDS up to rlds_cap32(W) positions.

__device__ __forceinline__ int pair_q(const DevState& S, int p) { return p / S.nr; }
__device__ __forceinline__ int pair_r(const DevState& S, int p) { return p % S.nr; }

// Next item of a device work queue for the calling wave (lane 0 takes it
// from the queue's counter; every lane returns it).
__device__ __forceinline__ int wq_next_wave(const WorkQ& wq) {
    int it = 0;
    if (lane_id() == 0) it = atomicAdd(wq.head, 1);
    return __shfl(it, 0, GTA_WAVE);
}
// The same for a whole workgroup (thread 0 takes it, LDS broadcast).
__device__ __forceinline__ int wq_next_block(const WorkQ& wq, int* slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(wq.head, 1);
    __syncthreads();
    const int it = *slot;
    __syncthreads();
    return it;
}

// convergence of the covariance record (CheckConvergenceRefined_Complete,
// MpStageBase.h:262-277): all 16 fields changed by less than 1e-5 relative
template <typename T>
__device__ __forceinline__ bool cc_converged(const T* a, T na, const T* b, T nb) {
    int n = 0;
#pragma unroll
    for (int f = 0; f < 15; f++) {
        T d = a[f] - b[f];
        T m = t_abs(a[f]) < t_abs(b[f]) ? t_abs(a[f]) : t_abs(b[f]);
        if (t_abs(d) < m * (T)RM_CONVEPSILON) n++;
    }
    {
        T d = na - nb;
        T m = t_abs(na) < t_abs(nb) ? t_abs(na) : t_abs(nb);
        if (t_abs(d) < m * (T)RM_CONVEPSILON) n++;
    }
    return n >= 16;
}

// ============================================================================
// secondary structure (protein C-alpha distance rule, MpSecStr.h:150-180)
__global__ void k_secstr(DevChunk C) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= C.nposs) return;
    int lo = 0, hi = C.nstrs - 1;
    while (lo < hi) {  // last structure whose offset is <= g
        const int mid = (lo + hi + 1) >> 1;
        if ((int)C.dst[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int s = lo;
    const int L = C.len[s];
    const int D = (int)C.dst[s];
    const int j = g - D;
    if (C.na[s]) {
        // nucleic acid: the pairing partner of residue j (CalcDistances_NASS,
        // sskna.cu:135-270; CPU twin MpSecStr.cpp:253-360; C3' bounds
        // 12.5-15 A, PM2DVectorFields.h:178-206): the complementary residue x
        // (|x - j| > 2) whose distance deviates least from 13.75 A, within
        // 1.25 A, the lowest x on exact ties; k_secstr_na marks the pairs
        const char a = C.rsd[D + j];
        const float xj = C.x[D + j], yj = C.y[D + j], zj = C.z[D + j];
        const float avg = 13.75f, dlt = 1.25f;
        float best = 9999.9f;
        int bx = -1;
        for (int x = 0; x < L; x++) {
            const char b = C.rsd[D + x];
            const bool pr = ((a == 'T' || a == 'U') && b == 'A') || (a == 'A' && (b == 'T' || b == 'U')) ||
                            (a == 'G' && (b == 'C' || b == 'U')) || ((a == 'C' || a == 'U') && b == 'G');
            if (!pr || abs(x - j) <= 2) continue;
            const float dev = fabsf(sqrtf(dist2(xj, yj, zj, C.x[D + x], C.y[D + x], C.z[D + x])) - avg);
            if (dlt < dev) continue;
            if (dev < best) { best = dev; bx = x; }
        }
        C.napos[D + j] = (bx >= 0 && best < 1.5f) ? bx : -1;
        return;
    }
    char v = ' ';
    if (2 <= j && j < L - 2) {
        auto d = [&](int a, int b) {
            return sqrtf(dist2(C.x[D + a], C.y[D + a], C.z[D + a], C.x[D + b], C.y[D + b], C.z[D + b]));
        };
        float d13 = d(j - 2, j), d14 = d(j - 2, j + 1), d15 = d(j - 2, j + 2);
        float d24 = d(j - 1, j + 1), d25 = d(j - 1, j + 2), d35 = d(j, j + 2);
        float e = 2.1f;
        if (fabsf(d15 - 6.37f) < e && fabsf(d14 - 5.18f) < e && fabsf(d25 - 5.18f) < e &&
            fabsf(d13 - 5.45f) < e && fabsf(d24 - 5.45f) < e && fabsf(d35 - 5.45f) < e)
            v = 'h';
        else {
            e = 1.42f;
            if (fabsf(d15 - 13.0f) < e && fabsf(d14 - 10.4f) < e && fabsf(d25 - 10.4f) < e &&
                fabsf(d13 - 6.1f) < e && fabsf(d24 - 6.1f) < e && fabsf(d35 - 6.1f) < e)
                v = 'e';
            else if (d15 < 8.0f)
                v = 't';
        }
    }
    C.ss[D + j] = v;
}

// Nucleic-acid SS from the pairing partners (MpSecStr.cpp:148-200): mutual
// pairs j < x become '(' at j and ')' at x, every other position '.'.
__global__ void k_secstr_na(DevChunk C) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= C.nposs) return;
    int lo = 0, hi = C.nstrs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)C.dst[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int s = lo;
    if (!C.na[s]) return;
    const int D = (int)C.dst[s], L = C.len[s], j = g - D;
    const int p = C.napos[g];
    char v = '.';
    if (0 <= p && p < L && j < p && C.napos[D + p] == j) v = '(';
    else if (0 <= p && p < L && p < j && C.napos[D + p] == j) v = ')';
    C.ss[g] = v;
}

// ============================================================================
__global__ void k_init(DevChunk Q, DevChunk R, DevState S) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.nq * S.nr) return;
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    S.flags[p] = (Q.na[qi] == R.na[ri]) ? 0 : CONV_LOWTMSC;
    S.grand[p] = 0.f; S.best0[p] = 0.f; S.prevbest0[p] = 0.f;
    S.nalnposs[p] = 0; S.qrypos[p] = 0; S.rfnpos[p] = 0; S.sfragndx[p] = 0; S.sfragpos[p] = 0;
    for (int i = 0; i < 12; i++) {
        float v = (i == 0 || i == 4 || i == 8) ? 1.f : 0.f;
        S.tfm_ibest[p * 12 + i] = v;
        S.tfm_grand[p * 12 + i] = v;
    }
    for (int i = 0; i < 15; i++) S.out[p * 15 + i] = 0.f;
    S.out[p * 15 + 8] = 9999.9f;
}

// The reference filter after the stage-1 prescreen (MpReform.h:117-200,
// MpReform.cpp:28-124; CuBatch.cu:558-563): per block of `qgroup` consecutive
// queries (0: all queries), a reference flagged low-scoring for every query
// of the block leaves it (its pairs keep the flag), and every pair of a
// reference that stays has its flags reset to the molecule-type check.  One
// thread per (block, reference).
__global__ void k_keep_refs(DevChunk Q, DevChunk R, DevState S, int qgroup) {
    const int nb = (S.nq + qgroup - 1) / qgroup;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nb * S.nr) return;
    const int b = (int)(t / S.nr), ri = (int)(t % S.nr);
    const int q0 = b * qgroup, q1 = min(S.nq, q0 + qgroup);
    bool stays = false;
    for (int qi = q0; qi < q1; qi++) stays |= !(S.flags[qi * S.nr + ri] & CONV_LOWTMSC);
    if (!stays) return;
    for (int qi = q0; qi < q1; qi++) S.flags[qi * S.nr + ri] = (Q.na[qi] == R.na[ri]) ? 0 : CONV_LOWTMSC;
}

// ============================================================================
// gapless passes over structure coordinates (stage 1), in slot groups
// One structure's coordinate planes from a workgroup-uniform base: position
// i is read at a 32-bit byte offset from the base (kept in scalar registers),
// which compiles to a global load with a scalar base and a 32-bit vector
// offset instead of per-lane 64-bit address arithmetic.
struct Coords {
    const float* x;
    const float* y;
    const float* z;
    __device__ __forceinline__ static float at(const float* b, int i) {
        return *(const float*)((const char*)b + ((unsigned)i << 2));
    }
    __device__ __forceinline__ void load(int i, float& px, float& py, float& pz) const {
        px = at(x, i); py = at(y, i); pz = at(z, i);
    }
};

__device__ __forceinline__ void copy12(float* d, const float* s) {
#pragma unroll
    for (int i = 0; i < 12; i++) d[i] = s[i];
}

// best slot of a pair: highest score, lowest slot index on ties
struct SlotBest {
    float best;
    int slot;
    float tfm[12];
    float gbest;
    int ndx, pos;
    __device__ __forceinline__ void clear() { best = 0.f; slot = 0x7fffffff; gbest = 0.f; ndx = 0; pos = 0; }
    __device__ __forceinline__ bool beats(float b, int s) const { return b > best || (b == best && s < slot); }
};
constexpr int NGROUPS = RWAVES * NGW;   // slot groups per workgroup
// The seed search runs narrower groups (more seeds per wave): its seed
// diagonals are many and short-lived, so each 3x3 solve, computed once per
// group, weighs more than in the refinement kernels.  Measured on the
// 16 x 50,000 workload: 16 lanes 0.623 s, 8 lanes 0.449 s, 4 lanes 0.378 s,
// 2 lanes 0.372 s, 1 lane 0.392 s.
#ifndef GTA_FINDFRAG_LANES
#define GTA_FINDFRAG_LANES 4
#endif
static_assert(GTA_FINDFRAG_LANES == 1 || GTA_FINDFRAG_LANES == 2 || GTA_FINDFRAG_LANES == 4 ||
                  GTA_FINDFRAG_LANES == 8 || GTA_FINDFRAG_LANES == 16,
              "seed-search groups of 1, 2, 4, 8 or 16 lanes");
constexpr int FF_GL = GTA_FINDFRAG_LANES;

// the workgroup's best over its groups of G lanes (every thread returns it)
template <int G = GL, int NT = RTHREADS>
__device__ void group_best(const SlotBest& mine, SlotBest* lds, SlotBest& out) {
    const int g = threadIdx.x / G;
    if (glane<G>() == 0) lds[g] = mine;
    __syncthreads();
    out = lds[0];
    for (int i = 1; i < NT / G; i++)
        if (out.beats(lds[i].best, lds[i].slot)) out = lds[i];
}
// the same when each of the workgroup's NG groups already keeps its record in lds[group]
template <int NG = NGROUPS>
__device__ void group_best_lds(const SlotBest* lds, SlotBest& out) {
    __syncthreads();
    out = lds[0];
    for (int i = 1; i < NG; i++)
        if (out.beats(lds[i].best, lds[i].slot)) out = lds[i];
}

// covariance over n diagonal positions from (qp, rp); with thr, only pairs
// within thr (squared distance under tfm); returns the group's count
template <int P, int G = GL>
__device__ __forceinline__ int gcc_gapless(const Coords& q, const Coords& r, int qp, int rp, int n,
                                           const float* tfm, bool usethr, float thr,
                                           typename GPrec<P>::T* cc) {
    typename GPrec<P>::Acc acc;
    acc.clear();
    int cnt = 0;
    for (int k = glane<G>(); k < n; k += G) {
        const int a = qp + k, b = rp + k;
        float qx, qy, qz, rx, ry, rz;
        q.load(a, qx, qy, qz);
        r.load(b, rx, ry, rz);
        if (usethr && thr < tfm_dist2(tfm, qx, qy, qz, rx, ry, rz)) continue;
        acc.add(qx, qy, qz, rx, ry, rz);
        cnt++;
    }
    acc.template reduce<G>(cc);
    return gsum<G>(cnt);
}

// seed-diagonal scores (CalcScoresUnrl_Complete, MpStage1.h:700-790): all
// pairs scored; next threshold from the 3rd smallest distinct distance
template <int P, int G = GL>
__device__ __forceinline__ float gscore_seed(const Coords& q, const Coords& r, int qp, int rp, int n,
                                             const float* tfm, float d0, float d02, bool calc2, bool want_thr,
                                             float* thr) {
    GScore<P> sc;
    sc.clear();
    Min3d m3;
    m3.init();
    // distances are tracked only when some group of the wave wants a threshold
    const bool track = __any(want_thr);
    for (int k = glane<G>(); k < n; k += G) {
        const int a = qp + k, b = rp + k;
        float qx, qy, qz, rx, ry, rz;
        q.load(a, qx, qy, qz);
        r.load(b, rx, ry, rz);
        const float d = tfm_dist2(tfm, qx, qy, qz, rx, ry, rz);
        sc.add(pair_score(d02, d));
        if (track) m3.put(d);
    }
    const float s = sc.template reduce<G>();
    if (want_thr) {
        float min3 = gthird<G>(m3);
        const float d0s = GetD0s(d0);
        float d02s = d0s * d0s;
        if (calc2) d02s += 1.0f;
        if (CP_LARGEDST_cmp < min3 || min3 < d02s || n <= 3) min3 = d02s;
        else min3 = d02s + ceilf((min3 - d02s) * 2.0f) * 0.5f;
        *thr = min3;
    }
    return s;
}

// refinement scores (CalcScoresUnrlRefined_Complete / _DPRefined_Complete):
// pairs within d8 (when chk) scored; threshold from the multiset 3rd smallest
// distance, rounded up to a 0.5 A step above d0s-1
__device__ __forceinline__ float refined_threshold(float min3, float d0, int n) {
    const float d0s = GetD0s(d0) - 1.0f;
    const float d02s = d0s * d0s;
    if (CP_LARGEDST_cmp < min3 || min3 < d02s || n <= 3) return d02s;
    const float t = d0s + ceilf((sqrtf(min3) - d0s) * 2.0f) * 0.5f;
    return t * t;
}

template <int P>
__device__ __forceinline__ float gscore_gapless(const Coords& q, const Coords& r, int qp, int rp, int n,
                                                const float* tfm, float d0, float d02, float d82,
                                                bool want_thr, float* thr) {
    GScore<P> sc;
    sc.clear();
    Min3 m3;
    m3.init();
    const bool track = __any(want_thr);
    for (int k = glane(); k < n; k += GL) {
        const int a = qp + k, b = rp + k;
        float qx, qy, qz, rx, ry, rz;
        q.load(a, qx, qy, qz);
        r.load(b, rx, ry, rz);
        const float d = tfm_dist2(tfm, qx, qy, qz, rx, ry, rz);
        if (d <= d82) sc.add(pair_score(d02, d));
        if (track) m3.put(d);
    }
    const float s = sc.reduce();
    if (want_thr) *thr = refined_threshold(gthird(m3), d0, n);
    return s;
}

// gscore_gapless (no threshold update) fused with the next iteration's
// thresholded covariance over the same positions under the same transform
// (gcc_gapless with threshold thrcc): bitwise the results of the two passes.
template <int P>
__device__ __forceinline__ float gscore_cc_gapless(const Coords& q, const Coords& r, int qp, int rp, int n,
                                                   const float* tfm, float d02, float d82, float thrcc,
                                                   typename GPrec<P>::T* cc, int* ncc) {
    GScore<P> sc;
    sc.clear();
    typename GPrec<P>::Acc acc;
    acc.clear();
    int cnt = 0;
    for (int k = glane(); k < n; k += GL) {
        const int a = qp + k, b = rp + k;
        float qx, qy, qz, rx, ry, rz;
        q.load(a, qx, qy, qz);
        r.load(b, rx, ry, rz);
        const float d = tfm_dist2(tfm, qx, qy, qz, rx, ry, rz);
        if (d <= d82) sc.add(pair_score(d02, d));
        if (!(thrcc < d)) {
            acc.add(qx, qy, qz, rx, ry, rz);
            cnt++;
        }
    }
    const float s = sc.reduce();
    acc.reduce(cc);
    *ncc = gsum(cnt);
    return s;
}

// ============================================================================
// stage 1: best gapless diagonal (seed) per pair (FindGaplessAlignedFragment,
// MpStage1.cpp:267-400).  Slots are seed diagonals; each runs up to three
// rounds of covariance -> Kabsch -> score.
__device__ __forceinline__ void findfrag_epilogue(const DevState& S, int p, const SlotBest& b) {
    S.best0[p] = b.best;
    if (b.best > 0.f) { S.qrypos[p] = b.ndx; S.rfnpos[p] = b.pos; }
}

// Slots [lo, hi) of part `part` of nparts of the range [0, n).
__device__ __forceinline__ void part_range(int n0, int n1, int part, int nparts, int& lo, int& hi) {
    const int n = max(n1 - n0, 0), per = (n + nparts - 1) / nparts;
    lo = min(n0 + part * per, n1);
    hi = min(lo + per, n1);
}

template <int P>
__global__ __launch_bounds__(FTHREADS) void k_findfrag(DevChunk Q, DevChunk R, DevState S, Geom G,
                                                        int dinv_on, const int* plist, SlotBest* parts) {
    constexpr int FG = FF_GL;
    __shared__ SlotBest lds[FTHREADS / FG];
    __shared__ int ctr;
    const int p = plist ? plist[blockIdx.x] : (int)blockIdx.x;
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    const int qlen = Q.len[qi], rlen = R.len[ri];
    // workgroup-uniform structure offsets (scalar registers)
    const int qd = __builtin_amdgcn_readfirstlane(Q.dst[qi]), rd = __builtin_amdgcn_readfirstlane(R.dst[ri]);
    if (S.flags[p] & CONV_LOWTMSC) {
        if (threadIdx.x == 0 && !parts) S.best0[p] = 0.f;
        return;
    }
    typedef typename GPrec<P>::T T;
    const Coords q{Q.x + qd, Q.y + qd, Q.z + qd}, r{R.x + rd, R.y + rd, R.z + rd};
    const bool dinv = dinv_on && (qlen < rlen);
    const float d0 = GetD0(qlen, rlen), d02 = d0 * d0;
    const int prv = max(min(qlen, rlen) >> 1, 5);
    // the pair's seeds: qlen-qpos >= prv and rlen-rpos >= prv, i.e.
    // abspos = n1 + si*stepinit in [prv-rlen, qlen-prv]
    const int slo = max(0, (prv - rlen - G.n1 + G.stepinit - 1) / G.stepinit);
    const int shi = min(G.nseeds, (qlen - prv - G.n1) / G.stepinit + 1);
    int lo, hi;
    part_range(slo, shi, blockIdx.y, gridDim.y, lo, hi);
    if (threadIdx.x == 0) ctr = lo;
    __syncthreads();

    SlotBest mine;
    mine.clear();
    unsigned long long ab = 0, af = 0;
    bool need = true, live = false;
    int si = 0, qpos = 0, rpos = 0, n = 0, round = 0;
    float tfm[12], thr = CP_LARGEDST, best = 0.f;
    for (;;) {
        // groups without a slot take the next valid seed diagonal
        for (;;) {
            const int got = fetch_slot<FG>(need, &ctr);
            if (need) {
                if (got >= hi) {
                    need = false;
                } else {
                    si = got;
                    const int abspos = G.n1 + si * G.stepinit;
                    qpos = max(0, abspos);
                    rpos = max(-abspos, 0);
                    if (!(qlen <= qpos || rlen <= rpos) && !(qlen - qpos < prv || rlen - rpos < prv)) {
                        n = min(qlen - qpos, rlen - rpos);
                        round = 0;
                        thr = CP_LARGEDST;
                        best = 0.f;
                        need = false;
                        live = true;
                    }
                }
            }
            if (!__any(need)) break;
        }
        if (!__any(live)) break;
        if (live) {
            T cc[15];
            const int cnt = gcc_gapless<P, FG>(q, r, qpos, rpos, n, tfm, round > 0, thr, cc);
            ab += (round > 0 ? AB_EXT : AB_CC) * (unsigned long long)n;
            af += (round > 0 ? AF_EXT : AF_CC) * (unsigned long long)n;
            bool done = false;
            if (round > 0 && (cnt < 1 || cnt == n)) {
                done = true;
            } else {
                kabsch<T>(cc, (T)cnt, dinv, tfm);
                const float sc = gscore_seed<P, FG>(q, r, qpos, rpos, n, tfm, d0, d02, round > 0, round < 2, &thr);
                ab += AB_TM * (unsigned long long)n + AB_SLOT;
                af += AF_TM * (unsigned long long)n + AF_SLOT;
                if (best < sc) best = sc;
                if (++round == 3) done = true;
            }
            if (done) {
                if (best > 0.f && mine.beats(best, si)) {
                    mine.best = best; mine.slot = si; mine.ndx = qpos; mine.pos = rpos;
                }
                live = false;
                need = true;
            }
        }
    }
    if (glane<FG>() == 0 && ab) {
        atomicAdd(S.algbytes + KC_FINDFRAG, ab);
        atomicAdd(S.algflops + KC_FINDFRAG, af);
    }
    SlotBest b;
    group_best<FG, FTHREADS>(mine, lds, b);
    if (threadIdx.x == 0) {
        if (parts) parts[(size_t)blockIdx.x * gridDim.y + blockIdx.y] = b;
        else findfrag_epilogue(S, p, b);
    }
}

// ============================================================================
// stage 1: refinement of fragments of the best gapless alignment
// (RefineFragInitKernel, MpStage1.cpp:470-640).  Slots are (fragment position,
// fragment length); each iterates to convergence.
__device__ __forceinline__ void gapless_epilogue(const DevState& S, int p, const SlotBest& b) {
    const bool have = b.best > 0.f;
    S.best0[p] = have ? b.best : 0.f;
    if (have) {
        copy12(S.tfm_ibest + p * 12, b.tfm);
        if (S.grand[p] < b.best) { S.grand[p] = b.best; copy12(S.tfm_grand + p * 12, b.tfm); }
    }
}

template <int P>
__global__ __launch_bounds__(RTHREADS) void k_refine_gapless(DevChunk Q, DevChunk R, DevState S, Geom G,
                                                              int dinv_on, const int* plist, SlotBest* parts) {
    __shared__ SlotBest lds[NGROUPS];
    __shared__ float tbest[NGROUPS][12];
    __shared__ typename GPrec<P>::T tlast[NGROUPS][16];
    __shared__ typename GPrec<P>::T tnext[NGROUPS][16];   // next iteration's covariance from a fused pass
    __shared__ int ctr;
    const int p = plist ? plist[blockIdx.x] : (int)blockIdx.x;
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    const int qlen = Q.len[qi], rlen = R.len[ri];
    // workgroup-uniform structure offsets (scalar registers)
    const int qd = __builtin_amdgcn_readfirstlane(Q.dst[qi]), rd = __builtin_amdgcn_readfirstlane(R.dst[ri]);
    if (S.flags[p] & CONV_LOWTMSC) {
        if (threadIdx.x == 0 && !parts) S.best0[p] = 0.f;
        return;
    }
    typedef typename GPrec<P>::T T;
    const Coords q{Q.x + qd, Q.y + qd, Q.z + qd}, r{R.x + rd, R.y + rd, R.z + rd};
    const bool dinv = dinv_on && (qlen < rlen);
    const int qpos = S.qrypos[p], rpos = S.rfnpos[p];
    const float d0 = GetD0(qlen, rlen), d02 = d0 * d0, d82 = GetD82(qlen, rlen);
    const float d0s1 = GetD0s(d0) + 1.0f, d02s1 = d0s1 * d0s1;
    const bool anyslot = !(qlen <= qpos || rlen <= rpos);
    const int nall = anyslot ? min(qlen - qpos, rlen - rpos) : 0;
    int lo, hi;
    part_range(0, G.nlocsteps, blockIdx.y, gridDim.y, lo, hi);
    // slots at fragment positions from nall + SFRAGSTEP on fail the slot test
    // below for every fragment length >= 1: not fetched at all
    const int vslots = ((nall + SFRAGSTEP - 1) / SFRAGSTEP + 1) * NMAXSUBFRAGS;
    lo = min(lo, vslots);
    hi = min(hi, vslots);
    if (threadIdx.x == 0) ctr = lo;
    __syncthreads();

    // per group: its best slot and the current slot's best transform, in LDS
    const int grp = threadIdx.x / GL;
    SlotBest& mine = lds[grp];
    float* tfmBest = tbest[grp];
    if (glane() == 0) mine.clear();
    unsigned long long ab = 0, af = 0;
    bool need = anyslot, live = false;
    int si = 0, cit = 0, fqp = 0, frp = 0, fl = 0;
    float tfm[12], thr = CP_LARGEDST, best = 0.f;
    T* last = tlast[grp];       // previous covariance record (15 sums, count)
    T* next = tnext[grp];       // covariance of the next iteration, when fused
    bool havenext = false;
    for (;;) {
        for (;;) {
            const int got = fetch_slot(need, &ctr);
            if (need) {
                if (got >= hi) {
                    need = false;
                } else {
                    si = got;
                    const int sfragfct = si / NMAXSUBFRAGS, sfragndx = si - sfragfct * NMAXSUBFRAGS;
                    const int sfragpos = sfragfct * SFRAGSTEP;
                    fl = GetFragLength(qlen, rlen, qpos, rpos, sfragndx);
                    if (fl >= 1 && !(qlen + SFRAGSTEP <= qpos + sfragpos + fl ||
                                     rlen + SFRAGSTEP <= rpos + sfragpos + fl)) {
                        int ql = qlen, rl = rlen;
                        fqp = qpos + sfragpos;
                        frp = rpos + sfragpos;
                        UpdateLengths(ql, rl, fqp, frp, fl);
                        cit = 0;
                        thr = CP_LARGEDST;
                        best = 0.f;
                        need = false;
                        live = true;
                    }
                }
            }
            if (!__any(need)) break;
        }
        if (!__any(live)) break;
        if (live) {
            T cc[15];
            T ncc;
            bool done = false;
            if (cit == 0) {
                ncc = (T)gcc_gapless<P>(q, r, fqp, frp, fl, tfm, false, 0.f, cc);
                ab += AB_CC * (unsigned long long)fl;
                af += AF_CC * (unsigned long long)fl;
            } else {
                if (havenext) {
                    // computed by the previous iteration's fused score pass
#pragma unroll
                    for (int f = 0; f < 15; f++) cc[f] = next[f];
                    ncc = next[15];
                    havenext = false;
                } else {
                    ncc = (T)gcc_gapless<P>(q, r, qpos, rpos, nall, tfm, true, (cit < 2) ? thr : d02s1, cc);
                }
                ab += AB_EXT * (unsigned long long)nall;
                af += AF_EXT * (unsigned long long)nall;
                if (cc_converged<T>(cc, ncc, last, last[15])) done = true;
            }
            if (!done && ncc < (T)1) done = true;
            if (!done) {
                if (glane() == 0) {
                    for (int f = 0; f < 15; f++) last[f] = cc[f];
                    last[15] = ncc;
                }
                kabsch<T>(cc, ncc, dinv, tfm);
                float sc;
                if (cit >= 1 && cit + 1 < NMAXCONVIT + 2) {
                    // the next iteration's covariance takes the fixed threshold
                    // d02s1 under this transform: one pass for both
                    T cn[15];
                    int nn;
                    sc = gscore_cc_gapless<P>(q, r, qpos, rpos, nall, tfm, d02, d82, d02s1, cn, &nn);
                    if (glane() == 0) {
                        for (int f = 0; f < 15; f++) next[f] = cn[f];
                        next[15] = (T)nn;
                    }
                    havenext = true;
                } else {
                    sc = gscore_gapless<P>(q, r, qpos, rpos, nall, tfm, d0, d02, d82, cit < 1, &thr);
                }
                ab += AB_TM * (unsigned long long)nall + AB_SLOT;
                af += AF_TM * (unsigned long long)nall + AF_SLOT;
                if (best < sc) {
                    best = sc;
                    if (glane() == 0) copy12(tfmBest, tfm);
                }
                if (++cit == NMAXCONVIT + 2) done = true;
            }
            if (done) {
                if (best > 0.f && mine.beats(best, si) && glane() == 0) {
                    mine.best = best; mine.slot = si; copy12(mine.tfm, tfmBest);
                }
                live = false;
                need = true;
            }
        }
    }
    if (glane() == 0 && ab) {
        atomicAdd(S.algbytes + KC_REFINE_GAPLESS, ab);
        atomicAdd(S.algflops + KC_REFINE_GAPLESS, af);
    }
    SlotBest b;
    group_best_lds(lds, b);
    if (threadIdx.x == 0) {
        if (parts) parts[(size_t)blockIdx.x * gridDim.y + blockIdx.y] = b;
        else gapless_epilogue(S, p, b);
    }
}

// ============================================================================
// DP fill with backtrace and extraction of matched pairs, one wave per pair.
//
// The Lq x Lr matrix is swept in strips of 64 query rows; lane l owns row
// 64*s + l and at step t computes column t - l, so every step is one
// anti-diagonal of the strip.  The up and diagonal predecessors come from lane
// l-1 through a cross-lane shift; lane 0 takes the previous strip's last row
// (staged with the reference columns) and lane 63's values leave through a
// second shift into the row the next strip reads.  Every lane packs its own
// 2-bit backtrace codes of 16 steps into one word: 64 words, one coalesced
// 256-byte store.  The walk from the bottom-right cell reads runs of 64 cells
// along the current direction in parallel.
enum { DPM_IBEST = 0, DPM_GRAND = 1, DPM_CONSTRAINED = 2 };

struct DpParams {
    int mode;
    int gap0;
    float gapopen;
    int usess;          // score kind (SC_*)
    float ssw;          // weight of a secondary-structure match
    int dpscan;
    float rrw;          // weight of the residue score (SC_SSRR)
};

// Cell score kinds of a DP fill: the TM-score-like distance score under the
// pair's superposition (SC_TM), with a bonus for matching secondary
// structure (SC_TMSS), or, with no superposition, secondary-structure match
// plus weighted residue similarity (SC_SSRR: ExecDPSSwBtck128xKernel with
// sequence scoring, MpDPHub.cpp:985-1230).
enum { SC_TM = 0, SC_TMSS = 1, SC_SSRR = 2 };

// SC_SSRR keeps gap extension free by storing the value of a cell reached
// diagonally shifted by DPSS_INIT (DPSSDEFINITSCOREVAL, mpdpbase.h:16): a
// neighbour below DPSS_INIT_cmp was reached diagonally, and leaving it
// sideways opens a gap.  Cells outside the matrix hold DPSS_INIT.
constexpr float DPSS_INIT = -1048576.0f;
constexpr float DPSS_INIT_cmp = -524288.0f;

// Residue scores of SC_SSRR in bits (_GONNET_SCORES_, alpha.h:36-60), with
// residue codes 'A'..'Z' -> 0..25 and any other character -> 26 (a zero row
// and column).  Uploaded from the host table with the length tables.
constexpr int NGT = 27;
__device__ float g_gonnet[NGT * NGT];

__device__ __forceinline__ int res_code(char c) { return ('A' <= c && c <= 'Z') ? c - 'A' : NGT - 1; }

// the residue scores times the weight, into LDS (the fills read them per cell)
__device__ __forceinline__ void load_gtab(float* gtab, float rrw) {
    for (int k = threadIdx.x; k < NGT * NGT; k += blockDim.x) gtab[k] = __fmul_rn(g_gonnet[k], rrw);
    __syncthreads();
}

// Backtrace word of lane l in block (s, tb): the codes of row 64s+l at the
// strip's steps 16tb..16tb+15, step t at bits 2(t&15); cell (i, j) is row i
// at step t = j + (i & 63).
__device__ __forceinline__ unsigned read_code(const unsigned* bt, int nblk, int i, int j) {
    const int s = i >> 6, l = i & 63, t = j + l;
    const unsigned w = bt[((size_t)(s * nblk + (t >> 4)) << 6) + l];
    return (w >> ((t & 15) << 1)) & 3u;
}

// whole-wave shifts by one lane (DPP wave_shr:1 and wave_shl:1): lane l
// receives lane l-1's (l+1's) value, lane 0 (63) receives `fill`
__device__ __forceinline__ float wave_shr1(float v, float fill) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                                 __builtin_bit_cast(int, v), 0x138, 0xf, 0xf,
                                                                 false));
}
__device__ __forceinline__ float wave_shl1(float v, float fill) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                                 __builtin_bit_cast(int, v), 0x130, 0xf, 0xf,
                                                                 false));
}
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Per-pair parameters of a DP fill.
struct DpPair {
    int p, qi, ri, Lq, Lr, qd, rd, flags;
    float tfm[12];
    float d02;
};

// false when the pair takes no DP in this mode (its matched count is set)
__device__ __forceinline__ bool dp_setup(const DevChunk& Q, const DevChunk& R, const DevState& S,
                                         const DpParams& prm, int p, DpPair& P) {
    // one pair per wave: its indices and lengths are wave-uniform (scalar
    // registers, scalar loop control in the strip sweep)
    p = __builtin_amdgcn_readfirstlane(p);
    P.p = p;
    P.qi = pair_q(S, p);
    P.ri = pair_r(S, p);
    P.Lq = __builtin_amdgcn_readfirstlane(Q.len[P.qi]);
    P.Lr = __builtin_amdgcn_readfirstlane(R.len[P.ri]);
    P.qd = __builtin_amdgcn_readfirstlane(Q.dst[P.qi]);
    P.rd = __builtin_amdgcn_readfirstlane(R.dst[P.ri]);
    P.flags = __builtin_amdgcn_readfirstlane(S.flags[p]);
    if (prm.mode == DPM_IBEST) {
        if (P.flags & (CONV_LOWTMSC | CONV_SCOREDP)) {
            if (threadIdx.x == 0) S.nalnposs[p] = 0;
            return false;
        }
    } else if (P.flags & CONV_LOWTMSC) {
        if (prm.mode == DPM_GRAND && threadIdx.x == 0) S.nalnposs[p] = 0;
        return false;
    }
    const float* tsrc = (prm.mode == DPM_IBEST ? S.tfm_ibest : S.tfm_grand) + (size_t)p * 12;
#pragma unroll
    for (int i = 0; i < 12; i++) P.tfm[i] = tsrc[i];
    float d0 = GetD0(P.Lq, P.Lr);
    if (prm.dpscan) d0 += 1.5f;
    P.d02 = d0 * d0;
    return true;
}

// Row handoff between the strips of one wave: nothing to wait for.
struct NoSync {
    __device__ __forceinline__ bool wait(int, int) { return true; }
    __device__ __forceinline__ void publish(int, int) {}
};

// LDS staging of one wave's reference columns: a ring of DPRING entries (two
// 64-column blocks) of {x, y, z, row-above value}, and the SS letters in a
// float ring of the same size.  Column c lives at entry c mod DPRING.
constexpr int DPRING = 128;

// One strip of 64 query rows of the DP (the strip sweep of every DP kernel).
// Lane l owns row 64s+l and at step t computes column t-l, so a step advances
// one anti-diagonal.  Each lane reads its column's entry from the LDS ring
// (prefetched a step ahead); the whole wave loads the next 64-column block
// coalesced from HBM a block ahead and writes it into the ring when the block
// it replaces has been read.  The row above the strip (rowin, written by the
// previous strip) comes in the same blocks and reaches lane 0 with the
// up-shift of the values; lane 63's values leave through a down-shift
// conveyor into rowout, stored and published a block at a time.  No divergent
// branches in the step loop.  SC is the cell score kind (SC_*); GAP0 fixes
// the gap mode at compile time; without BTCK no backtrace is written
// (score-only fills).  vend, when given, receives each lane's value at the
// strip's last column (row 64s+l).  SC_SSRR ring entries hold {residue code
// (int bits), SS letter, 0, row-above value} and read the weighted residue
// scores from gtab.  Returns false when a row wait gave up (Sync::wait).
// PUT: the strip stores its last row (rowout given), fixed at compile time.
// The steps where every lane's cell lies inside the matrix (a strip of 64 rows
// of the query, steps 63 to Lr-1) run a form without the bounds selects.
template <int SC, bool GAP0, bool BTCK, bool PUT, class Sync>
__device__ __forceinline__ bool dp_strip_impl(const DevChunk& Q, const DevChunk& R, const DevState& S,
                                              const DpParams& prm, const DpPair& P, int s, const float* rowin,
                                              float* rowout, unsigned* bt, float4* ring, float* ssring,
                                              Sync& sync, float* vend, const float* gtab) {
    constexpr bool USESS = SC == SC_TMSS;
    constexpr bool SSRR = SC == SC_SSRR;
    // value of the cells outside the matrix
    constexpr float OUT = SSRR ? DPSS_INIT : 0.0f;
    const int lane = lane_id();
    const int Lq = P.Lq, Lr = P.Lr;
    const float* Rx = R.x + P.rd;
    const float* Ry = R.y + P.rd;
    const float* Rz = R.z + P.rd;
    const char* Rs = R.ss + P.rd;
    const char* Rr = R.rsd + P.rd;
    const float gapopen = prm.gapopen, ssw = prm.ssw, d02 = P.d02;
    const int nsteps = Lr + 63;
    const int nblkw = (nsteps + 15) >> 4;
    const int i = (s << 6) + lane;
    const bool rowok = i < Lq;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    float qss = 0.f;
    int qbase = 0;      // SC_SSRR: the lane's row of gtab
    if (rowok) {
        if (SSRR) {
            qss = (float)Q.ss[P.qd + i];
            qbase = res_code(Q.rsd[P.qd + i]) * NGT;
        } else {
            qx = Q.x[P.qd + i]; qy = Q.y[P.qd + i]; qz = Q.z[P.qd + i];
            tfm_point(P.tfm, qx, qy, qz);
            if (USESS) qss = (float)Q.ss[P.qd + i];
        }
    }
    const bool qok = rowok && qx < DEFCOORD_QRY_cmp;
    const bool haveup = rowin != nullptr;
    constexpr bool putrow = PUT;
    if (haveup && !sync.wait(s - 1, min(128, Lr))) return false;
    // block 0 at entries 0..63, columns -64..-1 (entries 64..127) outside
    {
        const int c = lane;
        const bool ok = c < Lr;
        if (SSRR)
            ring[lane] = make_float4(__int_as_float(ok ? res_code(Rr[c]) : 0), ok ? (float)Rs[c] : 0.f, 0.f,
                                     (ok && haveup) ? rowin[c] : OUT);
        else
            ring[lane] = make_float4(ok ? Rx[c] : 0.f, ok ? Ry[c] : 0.f, ok ? Rz[c] : 0.f,
                                     (ok && haveup) ? rowin[c] : OUT);
        ring[64 + lane] = make_float4(0.f, 0.f, 0.f, OUT);
        if (USESS) {
            ssring[lane] = ok ? (float)Rs[c] : 0.f;
            ssring[64 + lane] = 0.f;
        }
    }
    // the next block waits in registers
    float nbx, nby, nbz, nbu, nbs = 0.f;
    {
        const int c2 = 64 + lane;
        const bool ok2 = c2 < Lr;
        if (SSRR) {
            nbx = __int_as_float(ok2 ? res_code(Rr[c2]) : 0); nby = ok2 ? (float)Rs[c2] : 0.f; nbz = 0.f;
        } else {
            nbx = ok2 ? Rx[c2] : 0.f; nby = ok2 ? Ry[c2] : 0.f; nbz = ok2 ? Rz[c2] : 0.f;
        }
        if (USESS) nbs = ok2 ? (float)Rs[c2] : 0.f;
        nbu = (ok2 && haveup) ? rowin[c2] : OUT;
    }
    int ao = ((-lane) & (DPRING - 1)) << 4;  // byte offset of column t - lane's ring entry
    float4 cur = *(const float4*)((const char*)ring + ao);
    float curs = USESS ? ssring[ao >> 4] : 0.f;
    float val = OUT, upprev = OUT, oblk = 0.f;
    unsigned word = 0;
    int j = -lane;
    // one step (anti-diagonal); false when a row wait gave up.  In an inner
    // step every lane's cell is in the matrix, so the bounds test (and the
    // column counter it reads) drops out.
    auto step = [&](const int t, auto kind) -> bool {
        // kind 0: edge step; 1: inner step; 2: inner step of a strip whose
        // query rows and reference columns all hold defined coordinates
        constexpr bool INNER = decltype(kind)::value >= 1;
        constexpr bool DEFINED = decltype(kind)::value == 2;
        const float4 rc = cur;
        const float rss = curs;
        if ((t & 63) == 63) {
            // block (t+1)/64 replaces columns t-127..t-64, all read by now
            const int e = (t + 1 + lane) & (DPRING - 1);
            ring[e] = make_float4(nbx, nby, nbz, nbu);
            if (USESS) ssring[e] = nbs;
            const int c2 = t + 65 + lane;
            const bool ok2 = c2 < Lr;
            if (SSRR) {
                nbx = __int_as_float(ok2 ? res_code(Rr[c2]) : 0); nby = ok2 ? (float)Rs[c2] : 0.f;
            } else {
                nbx = ok2 ? Rx[c2] : 0.f; nby = ok2 ? Ry[c2] : 0.f; nbz = ok2 ? Rz[c2] : 0.f;
            }
            if (USESS) nbs = ok2 ? (float)Rs[c2] : 0.f;
            if (haveup && t + 65 < Lr) {
                if (!sync.wait(s - 1, min(t + 129, Lr))) return false;
            }
            nbu = (ok2 && haveup) ? rowin[c2] : OUT;
        }
        ao = (ao + 16) & (DPRING * 16 - 1);
        cur = *(const float4*)((const char*)ring + ao);
        if (USESS) curs = ssring[ao >> 4];

        const float up = wave_shr1(val, rc.w);
        const float diag = upprev;
        upprev = up;
        const bool valid = INNER || (rowok && (unsigned)j < (unsigned)Lr);
        float v, nv;
        unsigned bk = BT_DIAG;
        if (SSRR) {
            // (float)(SS match) * wss + residue score * wrr, then the shifted
            // recurrence (MpDPHub.cpp:1150-1185)
            float sc = __fmul_rn(qss == rc.y ? 1.0f : 0.0f, ssw);
            sc = __fadd_rn(sc, gtab[qbase + __float_as_int(rc.x)]);
            float d = diag;
            if (d < DPSS_INIT_cmp) d = __fsub_rn(d, DPSS_INIT);
            v = __fadd_rn(sc, d);
            float w = val;
            if (w < DPSS_INIT_cmp) w = __fsub_rn(__fadd_rn(gapopen, w), DPSS_INIT);
            if (v < w) { v = w; bk = BT_LEFT; }
            w = up;
            if (w < DPSS_INIT_cmp) w = __fsub_rn(__fadd_rn(gapopen, w), DPSS_INIT);
            if (v < w) { v = w; bk = BT_UP; }
            nv = (bk != BT_DIAG) ? v : __fadd_rn(v, DPSS_INIT);
        } else {
            float sc = pair_score(d02, dist2(qx, qy, qz, rc.x, rc.y, rc.z));
            if (!DEFINED) sc = (qok && DEFCOORD_RFN_cmp < rc.x) ? sc : 0.0f;
            if (USESS) sc = __fadd_rn(sc, (qss == rss) ? ssw : 0.0f);
            v = __fadd_rn(sc, GAP0 ? diag : fabsf(diag));
            float w = val;
            if (!GAP0 && w < 0.0f) w = __fsub_rn(gapopen, w);
            if (v < w) { v = w; bk = BT_LEFT; }
            w = up;
            if (!GAP0 && w < 0.0f) w = __fsub_rn(gapopen, w);
            if (v < w) { v = w; bk = BT_UP; }
            nv = (GAP0 || bk != BT_DIAG) ? v : -v;
        }
        // cells left of the strip's columns keep 0, cells past them the last value
        val = valid ? nv : val;
        const unsigned code = valid ? bk : (unsigned)BT_STOP;
        if (!INNER) j++;
        if (putrow) {
            // lane l now holds column t-126+l of row 64s+63
            oblk = wave_shl1(oblk, val);
            const int jo = t - 63;
            if (jo >= 0 && ((jo & 63) == 63 || jo == Lr - 1)) {
                if (lane >= 63 - (jo & 63)) rowout[jo - 63 + lane] = oblk;
                sync.publish(s, jo + 1);
            }
        }
        if (BTCK) {
            word |= code << ((t & 15) << 1);
            if ((t & 15) == 15 || t == nsteps - 1) {
                bt[((size_t)(s * nblkw + (t >> 4)) << 6) + lane] = word;
                word = 0;
            }
        }
        return true;
    };
    // two steps per iteration, so that the prefetched ring entry and the
    // carried values alternate registers instead of being copied every step
    const std::integral_constant<int, 0> edge;
    const std::integral_constant<int, 1> inner;
    const std::integral_constant<int, 2> inner_defined;
    // inner steps: a full strip, steps 64 .. (Lr & ~1) - 1 (even bounds keep the pairs)
    const bool full = (s + 1) << 6 <= Lq;
    const int tin = full ? 64 : nsteps, tout = full ? max(64, Lr & ~1) : nsteps;
    // a full strip whose 64 query rows and Lr reference columns are all
    // defined scores every inner cell without the undefined-coordinate select
    // (the select would keep the score in every such cell).  The check reads
    // the reference's x once per strip (independent loads, no early exit);
    // references longer than 4,096 residues, whose strips the device-wide
    // pipeline runs back to back, keep the select instead.
    bool defined = false;
    if (!SSRR && full && Lr <= 4096) {
        bool d = qok;
#pragma unroll 4
        for (int c = lane; c < Lr; c += GTA_WAVE) d &= DEFCOORD_RFN_cmp < Rx[c];
        defined = __all(d);
    }
    int t = 0;
    for (; t + 1 < min(tin, nsteps); t += 2) {
        if (!step(t, edge)) return false;
        if (!step(t + 1, edge)) return false;
    }
    if (t < tout) {
        if (defined) {
            for (; t + 1 < tout; t += 2) {
                if (!step(t, inner_defined)) return false;
                if (!step(t + 1, inner_defined)) return false;
            }
        } else {
            for (; t + 1 < tout; t += 2) {
                if (!step(t, inner)) return false;
                if (!step(t + 1, inner)) return false;
            }
        }
        j = t - lane;   // the column counter as the edge steps would have left it
    }
    for (; t + 1 < nsteps; t += 2) {
        if (!step(t, edge)) return false;
        if (!step(t + 1, edge)) return false;
    }
    if (t < nsteps && !step(t, edge)) return false;
    if (vend) *vend = val;
    return true;
}
template <int SC, bool GAP0, bool BTCK = true, class Sync>
__device__ __forceinline__ bool dp_strip(const DevChunk& Q, const DevChunk& R, const DevState& S,
                                         const DpParams& prm, const DpPair& P, int s, const float* rowin,
                                         float* rowout, unsigned* bt, float4* ring, float* ssring, Sync& sync,
                                         float* vend = nullptr, const float* gtab = nullptr) {
    if (rowout)
        return dp_strip_impl<SC, GAP0, BTCK, true>(Q, R, S, prm, P, s, rowin, rowout, bt, ring, ssring, sync,
                                                   vend, gtab);
    return dp_strip_impl<SC, GAP0, BTCK, false>(Q, R, S, prm, P, s, rowin, rowout, bt, ring, ssring, sync, vend,
                                                gtab);
}

// Backtrace walk from the bottom-right cell (BtckToMatched, btck2match.cuh:46-52;
// ConstrainedBtckToMatched keeps pairs within d8 and their d2).  Vectorised:
// the 64 lanes read 64 successive cells of the current direction and a ballot
// finds the run length.  Matched pairs are stored in walk (reverse) order.
__device__ void dp_walk(const DevChunk& Q, const DevChunk& R, const DevState& S, const DpParams& prm,
                        const DpPair& P, const unsigned* bt) {
    const int lane = lane_id();
    const int p = P.p, Lq = P.Lq, Lr = P.Lr, qd = P.qd, rd = P.rd;
    if (prm.mode == DPM_GRAND && (P.flags & CONV_SCOREDP)) {
        if (lane == 0) S.nalnposs[p] = 0;
        return;
    }
    const int nblk = (Lr + 63 + 15) >> 4;
    const size_t base = (size_t)P.qi * (size_t)R.nposs + (size_t)rd;
    float* aqx = S.aln + base;
    float* aqy = aqx + S.alnplane;
    float* aqz = aqy + S.alnplane;
    float* arx = aqz + S.alnplane;
    float* ary = arx + S.alnplane;
    float* arz = ary + S.alnplane;
    float* ad2 = arz + S.alnplane;
    float* aqp = ad2 + S.alnplane;
    float* arp = aqp + S.alnplane;
    const bool constrained = prm.mode == DPM_CONSTRAINED;
    const float d82 = GetD82(Lq, Lr);
    int y = Lq - 1, x = Lr - 1, k = 0;
    unsigned dir = BT_DIAG;
    while (y >= 0 && x >= 0) {
        const int dy = (dir == BT_LEFT) ? 0 : 1, dx = (dir == BT_UP) ? 0 : 1;
        const int yy = y - lane * dy, xx = x - lane * dx;
        const bool inr = yy >= 0 && xx >= 0;
        const unsigned code = inr ? read_code(bt, nblk, yy, xx) : (unsigned)BT_STOP;
        const unsigned long long brk = __ballot(!(inr && code == dir));
        const int run = brk ? (int)__builtin_ctzll(brk) : GTA_WAVE;
        if (dir == BT_DIAG) {
            const bool take = lane < run;
            float qx = 0, qy = 0, qz = 0, rx = 0, ry = 0, rz = 0, d2 = 0;
            bool keep = take;
            if (take) {
                qx = Q.x[qd + yy]; qy = Q.y[qd + yy]; qz = Q.z[qd + yy];
                rx = R.x[rd + xx]; ry = R.y[rd + xx]; rz = R.z[rd + xx];
                if (constrained) {
                    d2 = tfm_dist2(P.tfm, qx, qy, qz, rx, ry, rz);
                    keep = d2 <= d82;
                }
            }
            const unsigned long long km = __ballot(keep);
            const int before = __popcll(km & ((1ull << lane) - 1ull));
            if (keep) {
                const int o = k + before;
                aqx[o] = qx; aqy[o] = qy; aqz[o] = qz;
                arx[o] = rx; ary[o] = ry; arz[o] = rz;
                aqp[o] = (float)yy; arp[o] = (float)xx;
                if (constrained) ad2[o] = d2;
            }
            k += __popcll(km);
        }
        y -= run * dy;
        x -= run * dx;
        if (run < GTA_WAVE) {
            const unsigned c = __shfl(code, run, GTA_WAVE);
            const int in = __shfl((int)inr, run, GTA_WAVE);
            if (!in || c == BT_STOP) break;
            dir = c;
        }
    }
    if (lane == 0) S.nalnposs[p] = k;
}

__device__ __forceinline__ void dp_count(const DevState& S, const DpPair& P) {
    if (threadIdx.x == 0) {
        const unsigned long long cells = (unsigned long long)P.Lq * (unsigned long long)P.Lr;
        atomicAdd(S.cells, cells);
        atomicAdd(S.algbytes + KC_DP, cells / 4 + 12ull * (unsigned long long)(P.Lq + P.Lr));
        atomicAdd(S.algflops + KC_DP, AF_CELL * cells);
        // the strip sweep's lane-steps: 64-row strips x 16-step blocks over Lr + 63 skewed steps
        atomicAdd(S.cells + 1, (unsigned long long)((P.Lq + 63) >> 6) *
                                   (unsigned long long)((P.Lr + 63 + 15) >> 4) * 1024ull);
    }
}

// DP fill with backtrace and the walk, one wave per pair (plist, when given,
// maps blocks to pairs).  The strips run in order; each reads the row the
// previous one wrote to the pair's row in HBM.
template <int SC, bool GAP0>
__device__ void dp_pair(const DevChunk& Q, const DevChunk& R, const DevState& S, const DpParams& prm, int p,
                        float4* ring, float* ssring, const float* gtab) {
    DpPair P;
    if (!dp_setup(Q, R, S, prm, p, P)) return;
    float* row = S.rowbuf + (size_t)P.qi * (size_t)R.nposs + (size_t)P.rd;
    unsigned* bt = S.btck + S.btckoff[P.p];
    const int nstrips = (P.Lq + 63) >> 6;
    NoSync ns;
    for (int s = 0; s < nstrips; s++) {
        dp_strip<SC, GAP0>(Q, R, S, prm, P, s, s > 0 ? row : nullptr, s + 1 < nstrips ? row : nullptr, bt, ring,
                           ssring, ns, nullptr, gtab);
        // the next strip reads this strip's row back
        __threadfence_block();
    }
    dp_count(S, P);
    __threadfence();
    dp_walk(Q, R, S, prm, P, bt);
}

// One wave per pair: block b takes pair plist[b] (b without a list), or, with
// a work queue, the waves of a persistent grid take queued pairs until the
// queue is drained.
template <int SC, bool GAP0>
__global__ __launch_bounds__(GTA_WAVE) __attribute__((amdgpu_waves_per_eu(SC == SC_TMSS ? 6 : 7))) void k_dp(
    DevChunk Q, DevChunk R, DevState S, DpParams prm,
                                                  const int* plist, WorkQ wq) {
    __shared__ float4 ring[DPRING];
    __shared__ float ssring[SC == SC_TMSS ? DPRING : 1];
    __shared__ float gtab[SC == SC_SSRR ? NGT * NGT : 1];
    if (SC == SC_SSRR) load_gtab(gtab, prm.rrw);
    if (wq.items) {
        const int n = *wq.count;
        for (int it = wq_next_wave(wq); it < n; it = wq_next_wave(wq))
            dp_pair<SC, GAP0>(Q, R, S, prm, wq.items[it], ring, ssring, gtab);
    } else {
        dp_pair<SC, GAP0>(Q, R, S, prm, plist ? plist[blockIdx.x] : (int)blockIdx.x, ring, ssring, gtab);
    }
}

// Score-only DP fills for several superpositions per pair (ExecDPScore128xKernel,
// MpDPHub.cpp:389-647, used by the frg3 superposition search to rank
// candidate transforms): for slot k of pair p the transform is
// tfms[(k*np + p)*12], and scores[k*np + p] receives the fill's value at the
// bottom-right cell (Lq-1, Lr-1), which the reference reads back as the
// slot's order-dependent score (MpStageFrg3.h:1186-1191).  Pairs flagged
// low-scoring and slots with a nonzero slotflags[k*np + p] are not filled and
// score 0.  One wave per pair sweeps its slots in turn through the pair's row
// buffer; no backtrace.
template <bool GAP0>
__global__ __launch_bounds__(GTA_WAVE) void k_dp_score(DevChunk Q, DevChunk R, DevState S, float gapopen,
                                                        const float* tfms, const int* slotflags, int nslots,
                                                        float* scores) {
    __shared__ float4 ring[DPRING];
    __shared__ float ssring[1];
    const int np = S.nq * S.nr;
    const int p = blockIdx.x;
    DpPair P;
    P.p = p;
    P.qi = pair_q(S, p);
    P.ri = pair_r(S, p);
    P.Lq = __builtin_amdgcn_readfirstlane(Q.len[P.qi]);
    P.Lr = __builtin_amdgcn_readfirstlane(R.len[P.ri]);
    P.qd = __builtin_amdgcn_readfirstlane(Q.dst[P.qi]);
    P.rd = __builtin_amdgcn_readfirstlane(R.dst[P.ri]);
    P.flags = __builtin_amdgcn_readfirstlane(S.flags[p]);
    const float d0 = GetD0(P.Lq, P.Lr);
    P.d02 = d0 * d0;
    const DpParams prm{DPM_GRAND, GAP0 ? 1 : 0, gapopen, SC_TM, 0.f, 0, 0.f};
    float* row = S.rowbuf + (size_t)P.qi * (size_t)R.nposs + (size_t)P.rd;
    const int nstrips = (P.Lq + 63) >> 6;
    const int lastlane = (P.Lq - 1) & 63;
    NoSync ns;
    for (int k = 0; k < nslots; k++) {
        const size_t o = (size_t)k * np + p;
        if ((P.flags & CONV_LOWTMSC) || (slotflags && slotflags[o])) {
            if (threadIdx.x == 0) scores[o] = 0.f;
            continue;
        }
#pragma unroll
        for (int i = 0; i < 12; i++) P.tfm[i] = tfms[o * 12 + i];
        float vend = 0.f;
        for (int s = 0; s < nstrips; s++) {
            dp_strip<SC_TM, GAP0, false>(Q, R, S, prm, P, s, s > 0 ? row : nullptr, s + 1 < nstrips ? row : nullptr,
                                         nullptr, ring, ssring, ns, &vend);
            __threadfence_block();
        }
        const float sc = readlane_f(vend, lastlane);
        if (threadIdx.x == 0) scores[o] = sc;
    }
}

// Row handoff between the waves of one workgroup: wave w sweeps strips
// w, w+W, ...; strip s writes its last row into ring row s mod W of the pair
// and publishes, in LDS, how many of its columns are stored.  The wave of
// strip s+1 waits (bounded spin) until the columns it loads next are there.
// The rows go through HBM with workgroup-scope release/acquire fences
// around the LDS progress counters.
template <int W>
struct LdsSync {
    int* strip;     // [W] strip each wave is on
    int* done;      // [W] columns of that strip's row stored
    int* fail;      // set when a wait gave up
    __device__ __forceinline__ bool wait(int sp, int cols) {
        const int pw = sp % W;
        for (long long it = 0;; it++) {
            const int cs = __hip_atomic_load(strip + pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const int cd = __hip_atomic_load(done + pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cs > sp || (cs == sp && cd >= cols)) break;
            if (it > (1ll << 24) || __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        return true;
    }
    __device__ __forceinline__ void publish(int s, int cols) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane_id() == 0) __hip_atomic_store(done + (s % W), cols, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ __forceinline__ void begin(int s) {
        if (lane_id() == 0) {
            __hip_atomic_store(done + (s % W), 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(strip + (s % W), s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
};

// DP of one large pair per workgroup of W waves (pairs from plist; ring rows
// of W x Lr floats per pair at ringoff[item]).  Same cells, codes and walk as
// k_dp; only the strips run concurrently, pipelined down the query.
template <int SC, bool GAP0, int W>
__global__ __launch_bounds__(W * GTA_WAVE) void k_dp_mw(DevChunk Q, DevChunk R, DevState S, DpParams prm,
                                                         const int* plist, float* ring,
                                                         const unsigned long long* ringoff, int* failflag) {
    constexpr bool USESS = SC == SC_TMSS;
    __shared__ int strip[W], done[W], fail;
    __shared__ float4 cring[W][DPRING];
    __shared__ float ssring[USESS ? W : 1][USESS ? DPRING : 1];
    __shared__ float gtab[SC == SC_SSRR ? NGT * NGT : 1];
    if (SC == SC_SSRR) load_gtab(gtab, prm.rrw);
    DpPair P;
    if (!dp_setup(Q, R, S, prm, plist[blockIdx.x], P)) return;
    const int w = threadIdx.x / GTA_WAVE;
    if (threadIdx.x < W) { strip[threadIdx.x] = -1; done[threadIdx.x] = 0; }
    if (threadIdx.x == 0) fail = 0;
    __syncthreads();
    float* rows = ring + ringoff[blockIdx.x];
    unsigned* bt = S.btck + S.btckoff[P.p];
    const int nstrips = (P.Lq + 63) >> 6;
    LdsSync<W> sy{strip, done, &fail};
    for (int s = w; s < nstrips; s += W) {
        sy.begin(s);
        const float* rin = s > 0 ? rows + (size_t)((s - 1) % W) * P.Lr : nullptr;
        float* rout = s + 1 < nstrips ? rows + (size_t)(s % W) * P.Lr : nullptr;
        if (!dp_strip<SC, GAP0>(Q, R, S, prm, P, s, rin, rout, bt, cring[w], ssring[USESS ? w : 0], sy, nullptr,
                                gtab))
            break;
        // the strip's row is complete
        sy.publish(s, P.Lr);
    }
    __syncthreads();
    if (fail) {
        if (threadIdx.x == 0) { atomicOr(failflag, 1); S.nalnposs[P.p] = 0; }
        return;
    }
    dp_count(S, P);
    if (w == 0) dp_walk(Q, R, S, prm, P, bt);
}

// Row handoff between workgroups anywhere on the device (k_dp_grid): strip s
// of a pair publishes in global memory how many columns of its last row are
// stored; the strip below waits (bounded spin) for the columns it loads next.
// Agent-scope release/acquire fences order the row stores and loads.
struct GridSync {
    int* done;      // per strip of the pair: columns of its row stored
    int* fail;      // set when a wait gave up; every other wait then gives up too
    __device__ __forceinline__ bool wait(int sp, int cols) {
        for (long long it = 0;; it++) {
            const int cd = __hip_atomic_load(done + sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cd >= cols) break;
            if (it > (1ll << 22) || __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                if (lane_id() == 0) __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return true;
    }
    __device__ __forceinline__ void publish(int s, int cols) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane_id() == 0) __hip_atomic_store(done + s, cols, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// DP fill of a few very large pairs by the whole device: one-wave workgroups
// take (pair, strip) items in increasing order from a global counter
// (sync[total]) and sweep them; strip s waits on strip s-1's published row
// (rows: nstrips x Lr floats per pair at rowoff[k]; strip0[k] is pair k's
// first item).  Items are handed out in order and a strip waits only on the
// strip before it, which a running workgroup already holds, so the lowest
// unfinished strip never waits and the fill always drains.  Same cells and
// codes as k_dp; k_dp_grid_walk then walks each pair.
template <int SC, bool GAP0>
__global__ __launch_bounds__(GTA_WAVE) void k_dp_grid(DevChunk Q, DevChunk R, DevState S, DpParams prm,
                                                       const int* plist, int npairs, const int* strip0,
                                                       int* sync, float* rows, const unsigned long long* rowoff,
                                                       int* failflag) {
    __shared__ float4 cring[DPRING];
    __shared__ float ssring[SC == SC_TMSS ? DPRING : 1];
    __shared__ float gtab[SC == SC_SSRR ? NGT * NGT : 1];
    if (SC == SC_SSRR) load_gtab(gtab, prm.rrw);
    const int total = strip0[npairs];
    int* next = sync + total;
    for (;;) {
        int item = 0;
        if (lane_id() == 0) item = atomicAdd(next, 1);
        item = __shfl(item, 0, GTA_WAVE);
        if (item >= total) break;
        int lo = 0, hi = npairs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (strip0[mid] <= item) lo = mid; else hi = mid - 1;
        }
        const int k = lo, s = item - strip0[k];
        DpPair P;
        if (!dp_setup(Q, R, S, prm, plist[k], P)) continue;
        const int nstrips = strip0[k + 1] - strip0[k];
        float* prow = rows + rowoff[k];
        unsigned* bt = S.btck + S.btckoff[P.p];
        GridSync sy{sync + strip0[k], failflag};
        const float* rin = s > 0 ? prow + (size_t)(s - 1) * P.Lr : nullptr;
        float* rout = s + 1 < nstrips ? prow + (size_t)s * P.Lr : nullptr;
        if (!dp_strip<SC, GAP0>(Q, R, S, prm, P, s, rin, rout, bt, cring, ssring, sy, nullptr, gtab)) break;
        sy.publish(s, P.Lr);
    }
}

// backtrace walk of the pairs k_dp_grid filled (one wave per pair)
__global__ __launch_bounds__(GTA_WAVE) void k_dp_grid_walk(DevChunk Q, DevChunk R, DevState S, DpParams prm,
                                                            const int* plist, const int* failflag) {
    DpPair P;
    if (!dp_setup(Q, R, S, prm, plist[blockIdx.x], P)) return;
    if (*failflag) {
        if (threadIdx.x == 0) S.nalnposs[P.p] = 0;
        return;
    }
    dp_count(S, P);
    dp_walk(Q, R, S, prm, P, S.btck + S.btckoff[P.p]);
}

// ============================================================================
// refinement over the DP alignment (reverse-ordered matched pairs)
// Refinement modes: DPRefine, the final stage's phases 1/2 and the production
// phases 1/2 (MpStageFin.h:337-460).  Phase 2 around phase 1's fragment:
// FIN2/PROD2 one fragment length at 2x128 positions (--refinement ONE/COARSE),
// FIN2F/PROD2F every fragment length at those positions (the full searches of
// --refinement FULLA and FULL/FULLA), PROD2L the sequential log search of
// --refinement LOG.
enum { RM_DPREFINE = 0, RM_FIN1 = 1, RM_FIN2 = 2, RM_PROD1 = 3, RM_PROD2 = 4, RM_FIN2F = 5, RM_PROD2F = 6,
       RM_PROD2L = 7 };
// the log search's steps: 32, 16, ..., 1 positions forward then backward,
// fragment lengths 1..5 at each (MpStageFin.cpp:1636-1663)
constexpr int LOG_NSTEPS = 6 * 2 * (NMAXSUBFRAGS - 1);

struct RefineParams {
    int dinv_on;
    // DPRefine bookkeeping (MpStage1.cpp:455-510); epi=0 disables it
    int epi;
    int dpi;
    int maxndpiters;
    float prescore;        // > 0 with prescreen at the last iteration
};

// The aligned pairs' coordinates, read one position at a time by the passes:
// six planes in HBM (Aln), or staged into LDS position-major, the six floats
// of position i at 6i (AlnL: three 8-byte LDS loads from one 32-bit address).
struct Aln {
    const float* qx; const float* qy; const float* qz;
    const float* rx; const float* ry; const float* rz;
    __device__ __forceinline__ void load(int i, float& x0, float& y0, float& z0, float& x1, float& y1,
                                         float& z1) const {
        x0 = qx[i]; y0 = qy[i]; z0 = qz[i]; x1 = rx[i]; y1 = ry[i]; z1 = rz[i];
    }
};
typedef float lds_f2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(3))) lds_f2* lds_cf2ptr;
struct AlnL {
    lds_cf2ptr p;
    __device__ __forceinline__ void load(int i, float& x0, float& y0, float& z0, float& x1, float& y1,
                                         float& z1) const {
        const lds_f2 a = p[3 * i], b = p[3 * i + 1], c = p[3 * i + 2];
        x0 = a.x; y0 = a.y; z0 = b.x; x1 = b.y; y1 = c.x; z1 = c.y;
    }
};

// forward alignment position f is stored at L-1-f
// covariance of the fragment [fragpos, fragpos+fl) of the alignment (with the
// sums of squares when sq is given, for the RMSD)
template <int P, class A>
__device__ __forceinline__ void gcc_dp_frag(const A& a, int L, int fragpos, int fl,
                                            typename GPrec<P>::T* cc, typename GPrec<P>::T* sq) {
    int ql = L, rl = L, qp = fragpos, rp = fragpos;
    UpdateLengths(ql, rl, qp, rp, fl);
    typedef typename GPrec<P>::T T;
    typename GPrec<P>::Acc acc;
    acc.clear();
    T s2[6] = {0, 0, 0, 0, 0, 0};
    for (int k = glane(); k < fl; k += GL) {
        const int idx = L - 1 - (rp + k);
        float qx, qy, qz, rx, ry, rz;
        a.load(idx, qx, qy, qz, rx, ry, rz);
        acc.add(qx, qy, qz, rx, ry, rz);
        if (sq) {
            // squares formed in T: exact for T = double
            s2[0] += (T)qx * (T)qx; s2[1] += (T)qy * (T)qy; s2[2] += (T)qz * (T)qz;
            s2[3] += (T)rx * (T)rx; s2[4] += (T)ry * (T)ry; s2[5] += (T)rz * (T)rz;
        }
    }
    acc.reduce(cc);
    if (sq) {
#pragma unroll
        for (int i = 0; i < 6; i++) sq[i] = gsum(s2[i]);
    }
}
// covariance of the aligned pairs within thr under tprev; returns the count
template <int P, class A>
__device__ __forceinline__ int gcc_dp_ext(const A& a, int L, const float* tprev, float thr,
                                          typename GPrec<P>::T* cc) {
    typename GPrec<P>::Acc acc;
    acc.clear();
    int cnt = 0;
    for (int f = glane(); f < L; f += GL) {
        const int idx = L - 1 - f;
        float qx, qy, qz, rx, ry, rz;
        a.load(idx, qx, qy, qz, rx, ry, rz);
        if (thr < tfm_dist2(tprev, qx, qy, qz, rx, ry, rz)) continue;
        acc.add(qx, qy, qz, rx, ry, rz);
        cnt++;
    }
    acc.reduce(cc);
    return gsum(cnt);
}
template <int P, class A>
__device__ __forceinline__ float gscore_dp(const A& a, int L, const float* tfm, float d0, float d02, float d82,
                                           bool chk, bool want_thr, float* thr) {
    GScore<P> sc;
    sc.clear();
    Min3 m3;
    m3.init();
    const bool track = __any(want_thr);
    for (int f = glane(); f < L; f += GL) {
        const int idx = L - 1 - f;
        float qx, qy, qz, rx, ry, rz;
        a.load(idx, qx, qy, qz, rx, ry, rz);
        const float d = tfm_dist2(tfm, qx, qy, qz, rx, ry, rz);
        if (!chk || d <= d82) sc.add(pair_score(d02, d));
        if (track) m3.put(d);
    }
    const float s = sc.reduce();
    if (want_thr) *thr = refined_threshold(gthird(m3), d0, L);
    return s;
}

// TM-score pass (as gscore_dp, no threshold update) fused with the next
// iteration's thresholded covariance pass under the same transform (as
// gcc_dp_ext with threshold thrcc): the positions, distances, accumulation
// order and reductions are those of the two separate passes, so both results
// are bitwise the same.  Returns the score; cc and *ncc receive the
// covariance and its count.
template <int P, class A>
__device__ __forceinline__ float gscore_cc_dp(const A& a, int L, const float* tfm, float d02, float d82, bool chk,
                                              float thrcc, typename GPrec<P>::T* cc, int* ncc) {
    GScore<P> sc;
    sc.clear();
    typename GPrec<P>::Acc acc;
    acc.clear();
    int cnt = 0;
    for (int f = glane(); f < L; f += GL) {
        const int idx = L - 1 - f;
        float qx, qy, qz, rx, ry, rz;
        a.load(idx, qx, qy, qz, rx, ry, rz);
        const float d = tfm_dist2(tfm, qx, qy, qz, rx, ry, rz);
        if (!chk || d <= d82) sc.add(pair_score(d02, d));
        if (!(thrcc < d)) {
            acc.add(qx, qy, qz, rx, ry, rz);
            cnt++;
        }
    }
    const float s = sc.reduce();
    acc.reduce(cc);
    *ncc = gsum(cnt);
    return s;
}

// Superposition refinement over the DP alignment, one workgroup per pair
// (FragmentBasedDPAlignmentRefinement, FinalFragmentBasedDPAlignmentRefinement
// Phase1/2, ProductionFragmentBasedDPAlignmentRefinementPhase1/2:
// MpStage1.cpp:455-510, MpStageFin.cpp:45-1470).  Slots are (fragment
// position, fragment length) along the alignment; each iterates covariance
// (thresholded after the first) -> Kabsch -> score to convergence.
template <int MODE>
__device__ void refine_dp_epilogue(const DevChunk& Q, const DevChunk& R, const DevState& S,
                                   const RefineParams& prm, int p, const SlotBest& b) {
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    const int qlen = Q.len[qi], rlen = R.len[ri];
    const bool prod = (MODE == RM_PROD1 || MODE == RM_PROD2 || MODE == RM_PROD2F);
    const bool have = b.best > 0.f;
    float* tg = S.tfm_grand + (size_t)p * 12;
    if (!prod) {
        S.best0[p] = have ? b.best : 0.f;
        if (have) copy12(S.tfm_ibest + (size_t)p * 12, b.tfm);
        if (MODE == RM_FIN1) { S.sfragndx[p] = have ? b.ndx : 0; S.sfragpos[p] = have ? b.pos : 0; }
        if (have && S.grand[p] < b.best) { S.grand[p] = b.best; copy12(tg, b.tfm); }
    } else {
        const float grand = (MODE == RM_PROD2 || MODE == RM_PROD2F) ? S.grand[p] : 0.f;
        if (have && grand < b.best) {
            S.grand[p] = b.best;
            if (MODE == RM_PROD1) { S.sfragndx[p] = b.ndx; S.sfragpos[p] = b.pos; }
            float sq = b.best, sr = b.gbest;
            if (rlen < qlen) { float t = sq; sq = sr; sr = t; }
            float* o = S.out + (size_t)p * 15;
            o[9] = sq / (float)qlen;
            o[10] = sr / (float)rlen;
            o[11] = GetD0fin(qlen, qlen, Q.na[qi]);
            o[12] = GetD0fin(rlen, rlen, R.na[ri]);
            copy12(tg, b.tfm);
        }
    }
    if (MODE == RM_DPREFINE && prm.epi) {
        int fl = S.flags[p];
        if (0 < prm.dpi && !(fl & (CONV_SCOREDP | CONV_LOWTMSC)) &&
            fabsf(S.best0[p] - S.prevbest0[p]) < SCORE_CONVEPSILON)
            fl |= CONV_SCOREDP;
        if (prm.dpi + 1 < prm.maxndpiters && !(fl & (CONV_SCOREDP | CONV_LOWTMSC))) S.prevbest0[p] = S.best0[p];
        if (prm.maxndpiters <= prm.dpi + 1) {
            if (0.0f < prm.prescore && S.grand[p] < prm.prescore * (float)min(qlen, rlen)) fl |= CONV_LOWTMSC;
            fl &= ~CONV_SCOREDP;
        }
        S.flags[p] = fl;
    }
}

// LDS of a refinement workgroup of W waves (W * NGW slot groups)
template <int P, int W>
struct RefineLds {
    SlotBest slds[W * NGW];
    float tbest[W * NGW][12];
    typename GPrec<P>::T tlast[W * NGW][16];
    typename GPrec<P>::T tnext[W * NGW][16];   // next iteration's covariance from a fused pass
    int ctr;
};

template <int P, int MODE, int W>
__device__ void refine_pair(const DevChunk& Q, const DevChunk& R, const DevState& S, const Geom& G,
                            const RefineParams& prm, int ldscap, int p, SlotBest* parts, float* lds_rf,
                            RefineLds<P, W>& rl) {
    SlotBest* slds = rl.slds;
    float (*tbest)[12] = rl.tbest;
    typename GPrec<P>::T (*tlast)[16] = rl.tlast;
    typename GPrec<P>::T (*tnext)[16] = rl.tnext;
    int& ctr = rl.ctr;
    typedef typename GPrec<P>::T T;
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    const int qlen = Q.len[qi], rlen = R.len[ri];
    const int flags = S.flags[p];
    const bool dinv = prm.dinv_on && (qlen < rlen);

    bool skip;
    if (MODE == RM_DPREFINE) skip = flags != 0;
    else skip = (flags & CONV_LOWTMSC) != 0;
    if (skip && MODE != RM_DPREFINE) return;

    const int L = skip ? 0 : S.nalnposs[p];
    const size_t base = (size_t)qi * (size_t)R.nposs + (size_t)R.dst[ri];
    const float* g = S.aln + base;
    const size_t pl = S.alnplane;
    const bool staged = !skip && L <= ldscap;
    if (staged) {
        for (int k = threadIdx.x; k < L; k += W * GTA_WAVE) {
            float* d = lds_rf + 6 * k;
            d[0] = g[k]; d[1] = g[pl + k]; d[2] = g[2 * pl + k];
            d[3] = g[3 * pl + k]; d[4] = g[4 * pl + k]; d[5] = g[5 * pl + k];
        }
    }
    __syncthreads();

    const bool prod = (MODE == RM_PROD1 || MODE == RM_PROD2 || MODE == RM_PROD2F || MODE == RM_PROD2L);
    const int rna = R.na[ri];
    const float d0 = prod ? GetD0fin(qlen, rlen, rna) : GetD0(qlen, rlen);
    const float d02 = d0 * d0, d82 = GetD82(qlen, rlen);
    const float d0s1 = GetD0s(d0) + 1.0f, d02s1 = d0s1 * d0s1;
    const int glen = max(qlen, rlen);
    const float g0 = GetD0fin(glen, glen, rna), g02 = g0 * g0;
    const bool chk = !prod;
    const float bestinit = (MODE == RM_DPREFINE) ? 0.f : -1.f;
    const bool phase2 = (MODE == RM_FIN2 || MODE == RM_PROD2);
    const bool full2 = (MODE == RM_FIN2F || MODE == RM_PROD2F);
    const bool logm = (MODE == RM_PROD2L);
    const int nslots = skip ? 0
                            : (phase2 ? G.maxnfragfcts
                                      : full2 ? G.maxnfragfcts * NMAXSUBFRAGS : logm ? LOG_NSTEPS : G.nlocsteps);
    const int base_ndx = (phase2 || full2) ? S.sfragndx[p] : 0;
    const int base_pos = (phase2 || full2 || logm) ? S.sfragpos[p] : 0;
    // log search: the running best (from the pair's grand best) and the
    // position the search has moved to, carried from step to step
    const float bestorg = logm ? S.grand[p] : 0.f;
    float lbest = bestorg, lbestprev = 0.f;
    int lpos = base_pos, lcenter = base_pos;
    int slo, shi;
    part_range(0, nslots, blockIdx.y, gridDim.y, slo, shi);
    if (!phase2 && !full2 && !logm) {
        // slots at fragment positions from L + SFRAGSTEP on fail the slot
        // test below for every fragment length >= 1: not fetched at all
        const int vslots = ((L + SFRAGSTEP - 1) / SFRAGSTEP + 1) * NMAXSUBFRAGS;
        slo = min(slo, vslots);
        shi = min(shi, vslots);
    }
    if (threadIdx.x == 0) ctr = slo;
    __syncthreads();

    // per group: its best slot and the current slot's best transform, in LDS
    const int grp = threadIdx.x / GL;
    SlotBest& mine = slds[grp];
    float* tfmBest = tbest[grp];
    if (glane() == 0) mine.clear();
    unsigned long long ab = 0, af = 0;
    // the log search's steps depend on each other: one slot group takes them in order
    bool need = slo < shi && (!logm || grp == 0), live = false;
    int si = 0, cit = 0, fl = 0, sfragndx = 0, sfragpos = 0;
    float tfm[12], thr = CP_LARGEDST, best = 0.f;
    T* last = tlast[grp];       // previous covariance record (15 sums, count)
    T* next = tnext[grp];       // covariance of the next iteration, when fused
    bool havenext = false;
    // the slot loop, instantiated for planes in LDS and for planes in HBM
    auto slots = [&](const auto& a) {
    for (;;) {
        for (;;) {
            const int got = fetch_slot(need, &ctr);
            if (need) {
                if (got >= shi) {
                    need = false;
                } else {
                    si = got;
                    int step;
                    if (phase2) {
                        sfragndx = base_ndx;
                        if (sfragndx == 0) sfragndx++;
                        sfragpos = base_pos + (si - (G.maxnfragfcts >> 1));
                        step = 1;
                    } else if (full2) {
                        const int sfragfct = si / NMAXSUBFRAGS;
                        sfragndx = si - sfragfct * NMAXSUBFRAGS;
                        sfragpos = base_pos + (sfragfct - (G.maxnfragfcts >> 1));
                        step = 1;
                    } else if (logm) {
                        const int sgnstep = ((si / (NMAXSUBFRAGS - 1)) & 1) ? -(32 >> (si / (2 * (NMAXSUBFRAGS - 1))))
                                                                          : (32 >> (si / (2 * (NMAXSUBFRAGS - 1))));
                        sfragndx = 1 + si % (NMAXSUBFRAGS - 1);
                        // the step's position is taken from the search's position
                        // when the step begins, before its fragment lengths run
                        if (sfragndx == 1) lcenter = lpos;
                        sfragpos = lcenter + sgnstep;
                        step = max(0, sgnstep);
                    } else {
                        const int sfragfct = si / NMAXSUBFRAGS;
                        sfragndx = si - sfragfct * NMAXSUBFRAGS;
                        sfragpos = sfragfct * SFRAGSTEP;
                        step = SFRAGSTEP;
                    }
                    fl = GetFragLength(L, L, 0, 0, sfragndx);
                    if (fl >= 1 && !(sfragpos < 0 || L + step <= sfragpos + fl)) {
                        cit = 0;
                        thr = CP_LARGEDST;
                        best = logm ? lbest : bestinit;
                        lbestprev = lbest;
#pragma unroll
                        for (int i = 0; i < 12; i++) tfm[i] = 0.f;
                        if (glane() == 0 && !logm)
                            for (int i = 0; i < 12; i++) tfmBest[i] = (i == 0 || i == 4 || i == 8) ? 1.f : 0.f;
                        need = false;
                        live = true;
                    }
                }
            }
            if (!__any(need)) break;
        }
        if (!__any(live)) break;
        if (live) {
            T cc[15];
            T ncc;
            bool done = false;
            if (cit == 0) {
                if (MODE == RM_PROD1 && si == 0 && P == 0) {
                    // the RMSD of the whole alignment comes with the first slot's
                    // first pass (the mixed form's comes from k_rmsd_fp64)
                    T sq[6];
                    gcc_dp_frag<P>(a, L, sfragpos, fl, cc, sq);
                    T ccr[15];
#pragma unroll
                    for (int f = 0; f < 15; f++) ccr[f] = cc[f];
                    const float rmsd = rmsd_e0<T>(ccr, sq, (T)fl);
                    if (glane() == 0) S.out[(size_t)p * 15 + 8] = rmsd;
                } else {
                    gcc_dp_frag<P>(a, L, sfragpos, fl, cc, nullptr);
                }
                ncc = (T)fl;
                ab += AB_CC * (unsigned long long)fl;
                af += AF_CC * (unsigned long long)fl;
            } else {
                if (havenext) {
                    // computed by the previous iteration's fused score pass
#pragma unroll
                    for (int f = 0; f < 15; f++) cc[f] = next[f];
                    ncc = next[15];
                    havenext = false;
                } else {
                    ncc = (T)gcc_dp_ext<P>(a, L, tfm, (cit < 2) ? thr : d02s1, cc);
                }
                ab += AB_EXT * (unsigned long long)L;
                af += AF_EXT * (unsigned long long)L;
                if (cc_converged<T>(cc, ncc, last, last[15])) done = true;
            }
            if (!done && ncc < (T)1) done = true;
            if (!done) {
                if (glane() == 0) {
                    for (int f = 0; f < 15; f++) last[f] = cc[f];
                    last[15] = ncc;
                }
                kabsch<T>(cc, ncc, dinv, tfm);
                float sc;
                if (cit >= 1 && cit + 1 < NMAXCONVIT + 2) {
                    // the next iteration's covariance takes the fixed threshold
                    // d02s1 under this transform: one pass for both
                    T cn[15];
                    int nn;
                    sc = gscore_cc_dp<P>(a, L, tfm, d02, d82, chk, d02s1, cn, &nn);
                    if (glane() == 0) {
                        for (int f = 0; f < 15; f++) next[f] = cn[f];
                        next[15] = (T)nn;
                    }
                    havenext = true;
                } else {
                    sc = gscore_dp<P>(a, L, tfm, d0, d02, d82, chk, cit < 1, &thr);
                }
                ab += AB_TM * (unsigned long long)L + AB_SLOT;
                af += AF_TM * (unsigned long long)L + AF_SLOT;
                if (best < sc) {
                    best = sc;
                    if (glane() == 0) copy12(tfmBest, tfm);
                }
                if (++cit == NMAXCONVIT + 2) done = true;
            }
            if (done && logm) {
                // the search moves to this position when it raised the best
                if (lbestprev < best) lpos = sfragpos;
                lbest = best;
                live = false;
                need = true;
            } else if (done) {
                float gbest = best;
                if (prod && qlen != rlen) {
                    float tb[12];
                    copy12(tb, tfmBest);
                    gbest = gscore_dp<P>(a, L, tb, g0, g02, d82, false, false, nullptr);
                    ab += AB_TM * (unsigned long long)L;
                    af += AF_TM * (unsigned long long)L;
                }
                if (best > 0.f && mine.beats(best, si) && glane() == 0) {
                    mine.best = best; mine.slot = si; copy12(mine.tfm, tfmBest);
                    mine.gbest = gbest; mine.ndx = sfragndx; mine.pos = sfragpos;
                }
                live = false;
                need = true;
            }
        }
    }
    if (logm) {
        // a result above the grand best goes directly to the output record and
        // transform (SaveBestQRScoresAndTM_Phase2_logsearch_Complete,
        // MpStageFin.h:777-801); every group computes, group 0 writes
        float tb[12];
        copy12(tb, tfmBest);
        float gbest = lbest;
        if (qlen != rlen) {
            gbest = gscore_dp<P>(a, L, tb, g0, g02, d82, false, false, nullptr);
            ab += AB_TM * (unsigned long long)L;
            af += AF_TM * (unsigned long long)L;
        }
        if (grp == 0 && glane() == 0 && bestorg < lbest) {
            float sq = lbest, sr = gbest;
            if (rlen < qlen) { float t = sq; sq = sr; sr = t; }
            float* o = S.out + (size_t)p * 15;
            o[9] = sq / (float)qlen;
            o[10] = sr / (float)rlen;
            copy12(S.tfm_grand + (size_t)p * 12, tb);
        }
    }
    };
    if (staged) {
        slots(AlnL{(lds_cf2ptr)lds_rf});
    } else {
        slots(Aln{g, g + pl, g + 2 * pl, g + 3 * pl, g + 4 * pl, g + 5 * pl});
    }
    if (glane() == 0 && ab) {
        atomicAdd(S.algbytes + (MODE == RM_DPREFINE ? KC_REFINE_DP : KC_REFINE_FIN), ab);
        atomicAdd(S.algflops + (MODE == RM_DPREFINE ? KC_REFINE_DP : KC_REFINE_FIN), af);
    }
    if (logm) return;
    SlotBest b;
    group_best_lds<W * NGW>(slds, b);
    if (threadIdx.x == 0) {
        if (parts) parts[(size_t)blockIdx.x * gridDim.y + blockIdx.y] = b;
        else refine_dp_epilogue<MODE>(Q, R, S, prm, p, b);
    }
}

// One workgroup of W waves per pair (plist[b], or b), each pair's slots split
// over gridDim.y workgroups when parts is given, or a persistent grid draining
// a work queue of pairs.
template <int P, int MODE, int W = RWAVES>
__global__ __launch_bounds__(W * GTA_WAVE) __attribute__((amdgpu_waves_per_eu(P == 0 ? 4 : 1))) void k_refine_dp(DevChunk Q, DevChunk R, DevState S, Geom G,
                                                         RefineParams prm, int ldscap, const int* plist,
                                                         SlotBest* parts, WorkQ wq) {
    extern __shared__ float lds_rf[];
    __shared__ RefineLds<P, W> rl;
    __shared__ int item;
    if (wq.items) {
        const int n = *wq.count;
        for (int it = wq_next_block(wq, &item); it < n; it = wq_next_block(wq, &item)) {
            refine_pair<P, MODE, W>(Q, R, S, G, prm, ldscap, wq.items[it], nullptr, lds_rf, rl);
            __syncthreads();
        }
    } else {
        refine_pair<P, MODE, W>(Q, R, S, G, prm, ldscap, plist ? plist[blockIdx.x] : (int)blockIdx.x, parts,
                                lds_rf, rl);
    }
}

// ============================================================================
// production alignment rows and statistics, one wave per pair
__device__ void prod_aln_pair(const DevChunk& Q, const DevChunk& R, const DevState& S, char* alns, int nodel,
                              int p) {
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    if (S.flags[p] & CONV_LOWTMSC) return;
    const int lane = lane_id();
    const int qlen = Q.len[qi];
    const int qd = Q.dst[qi], rd = R.dst[ri];
    const int m = S.nalnposs[p];
    const size_t base = (size_t)qi * (size_t)R.nposs + (size_t)rd;
    const float* ad2 = S.aln + base + 6 * S.alnplane;
    const float* aqp = ad2 + S.alnplane;
    const float* arp = aqp + S.alnplane;
    const size_t alnofff = ((size_t)qi * R.nposs + (size_t)S.nr * (size_t)(qd + qi)) * 5;
    const size_t dbalnlen = (size_t)R.nposs + (size_t)S.nr * (size_t)(qlen + 1);
    const size_t dbalnbeg = (size_t)rd + (size_t)ri * (size_t)(qlen + 1);
    char* row0 = alns + alnofff + dbalnbeg;
    auto put = [&](int col, char a, char mid, char b, char sa, char sb) {
        row0[col] = a; row0[dbalnlen + col] = mid; row0[2 * dbalnlen + col] = b;
        row0[3 * dbalnlen + col] = sa; row0[4 * dbalnlen + col] = sb;
    };
    int qp0 = 0, rp0 = 0, qpl = 0, rpl = 0;
    if (m > 0) {
        qp0 = (int)aqp[m - 1]; rp0 = (int)arp[m - 1];
        qpl = (int)aqp[0]; rpl = (int)arp[0];
    }
    int psts = 0, idts = 0;
    for (int k = lane; k < m; k += GTA_WAVE) {
        const int idx = m - 1 - k;
        const int qp = (int)aqp[idx], rp = (int)arp[idx];
        const float d2 = ad2[idx];
        const int col = nodel ? (qp - qp0) : (qp - qp0) + (rp - rp0) - k;
        const char qc = Q.rsd[qd + qp], tc = R.rsd[rd + rp];
        char mid = (d2 < EQUIV_D2) ? '+' : ' ';
        if (d2 < EQUIV_D2) psts++;
        if (qc != '-' && qc == tc) { idts++; mid = (mid == '+') ? qc : (char)(qc | 32); }
        put(col, qc, mid, tc, Q.ss[qd + qp], R.ss[rd + rp]);
        if (k + 1 < m) {
            const int qn = (int)aqp[idx - 1], rn = (int)arp[idx - 1];
            int c = col + 1;
            for (int q = qp + 1; q < qn; q++, c++) put(c, Q.rsd[qd + q], ' ', '-', Q.ss[qd + q], ' ');
            if (!nodel)
                for (int r = rp + 1; r < rn; r++, c++) put(c, '-', ' ', R.rsd[rd + r], ' ', R.ss[rd + r]);
        }
    }
    psts = wave_sum_i(psts);
    idts = wave_sum_i(idts);
    const int total = (m > 0) ? ((nodel ? (qpl - qp0) : (qpl - qp0) + (rpl - rp0) - (m - 1)) + 1) : 0;
    if (lane < 5) row0[(size_t)lane * dbalnlen + total] = 0;
    if (lane == 0) {
        float* o = S.out + (size_t)p * 15;
        unsigned bc = (m > 0) ? (((unsigned)(rp0 + 1) << 16) | (unsigned)(qp0 + 1)) : 0u;
        unsigned ec = (m > 0) ? (((unsigned)(rpl + 1) << 16) | (unsigned)(qpl + 1)) : 0u;
        o[2] = __uint_as_float(bc);
        o[3] = __uint_as_float(ec);
        o[4] = (float)total;
        o[5] = (float)psts;
        o[6] = (float)idts;
        o[7] = (float)(total - m);
    }
}

__global__ __launch_bounds__(GTA_WAVE) void k_prod_aln(DevChunk Q, DevChunk R, DevState S, char* alns,
                                                       int nodel, WorkQ wq) {
    if (wq.items) {
        const int n = *wq.count;
        for (int it = wq_next_wave(wq); it < n; it = wq_next_wave(wq)) prod_aln_pair(Q, R, S, alns, nodel, wq.items[it]);
    } else {
        prod_aln_pair(Q, R, S, alns, nodel, (int)blockIdx.x);
    }
}

// ============================================================================
// work lists of live pairs (no CONV_LOWTMSC), in pair order, for the batched
// form: list A for the one-wave DP (pairs of the multi-wave bucket left out,
// bigmask bit 0), list B for one-workgroup refinement (split pairs left out,
// bit 1).  Three passes: per-block counts, a scan, the scatter.
constexpr int LBLK = 1024;

// list membership of pair p: a = live and filled by the one-wave DP, b = live
// and refined by one workgroup, c = live (every pair that produces output)
__device__ __forceinline__ void live_flags(const DevState& S, const unsigned char* bigmask, int p, bool& a,
                                           bool& b, bool& c) {
    const int np = S.nq * S.nr;
    const bool live = p < np && !(S.flags[p] & CONV_LOWTMSC);
    const unsigned m = (bigmask && p < np) ? bigmask[p] : 0u;
    a = live && !(m & 1u);
    b = live && !(m & 2u);
    c = live;
}

__global__ __launch_bounds__(LBLK) void k_live_count(DevState S, const unsigned char* bigmask, int* cnt) {
    const int p = blockIdx.x * LBLK + threadIdx.x;
    bool a, b, c;
    live_flags(S, bigmask, p, a, b, c);
    const int ca = __syncthreads_count(a);
    const int cb = __syncthreads_count(b);
    const int cc = __syncthreads_count(c);
    if (threadIdx.x == 0) {
        cnt[3 * blockIdx.x] = ca;
        cnt[3 * blockIdx.x + 1] = cb;
        cnt[3 * blockIdx.x + 2] = cc;
    }
}

// exclusive scan of the per-block counts (one thread; nblk is np/1024) and
// the totals of the three lists
__global__ void k_live_scan(int* cnt, int nblk, int* totals) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int sa = 0, sb = 0, sc = 0;
    for (int i = 0; i < nblk; i++) {
        const int ca = cnt[3 * i], cb = cnt[3 * i + 1], cc = cnt[3 * i + 2];
        cnt[3 * i] = sa; cnt[3 * i + 1] = sb; cnt[3 * i + 2] = sc;
        sa += ca; sb += cb; sc += cc;
    }
    totals[0] = sa;
    totals[1] = sb;
    totals[2] = sc;
}

// listc may be null: the caller then drains lista for the output pass (the
// lists coincide when no pair is in the multi-wave DP bucket)
__global__ __launch_bounds__(LBLK) void k_live_scatter(DevState S, const unsigned char* bigmask, const int* off,
                                                       int* lista, int* listb, int* listc) {
    __shared__ int wa[LBLK / GTA_WAVE], wb[LBLK / GTA_WAVE], wc[LBLK / GTA_WAVE];
    const int p = blockIdx.x * LBLK + threadIdx.x;
    bool a, b, c;
    live_flags(S, bigmask, p, a, b, c);
    const unsigned long long ba = __ballot(a), bb = __ballot(b), bc = __ballot(c);
    const int w = threadIdx.x / GTA_WAVE, lane = lane_id();
    if (lane == 0) { wa[w] = __popcll(ba); wb[w] = __popcll(bb); wc[w] = __popcll(bc); }
    __syncthreads();
    int oa = off[3 * blockIdx.x], ob = off[3 * blockIdx.x + 1], oc = off[3 * blockIdx.x + 2];
    for (int i = 0; i < w; i++) { oa += wa[i]; ob += wb[i]; oc += wc[i]; }
    const unsigned long long below = (1ull << lane) - 1ull;
    if (a) lista[oa + __popcll(ba & below)] = p;
    if (b) listb[ob + __popcll(bb & below)] = p;
    if (c && listc) listc[oc + __popcll(bc & below)] = p;
}

// fills n ints with v (a kernel rather than hipMemsetAsync: the batched form
// replays the launch sequence as a HIP graph)
__global__ void k_fill_i32(int* d, int v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = v;
}

// ============================================================================
// Mixed precision: the RMSD of each pair's final alignment under its optimal
// superposition, Σ|Rq+t-r|²/n = (Sq + Sr - 2 λmax) / n, where λmax is the
// largest eigenvalue of Horn's 4x4 quaternion matrix of the centred
// covariance (Horn 1987, J. Opt. Soc. Am. A 4:629).  The sums are FP64 from
// exact products of the FP32 coordinates and λmax comes from cyclic Jacobi
// rotations, accurate to FP64 rounding of |N| even when eigenvalues nearly
// coincide; the E0 identity's cubic (rmsd_e0) loses half its digits there,
// i.e. on near-collinear or near-planar alignments.  Same pairs and positions
// as the E0 RMSD of the PROD1 refinement: the whole alignment of every pair
// that is not low-scoring.

// one Jacobi rotation annihilating a[I][J] (indices fixed at compile time so
// the matrix stays in registers)
template <int I, int J>
__device__ __forceinline__ void jacobi_rot(double (&a)[4][4]) {
    const double apq = a[I][J];
    if (apq == 0.0) return;
    const double th = (a[J][J] - a[I][I]) / (2.0 * apq);
    const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
    a[I][I] -= t * apq;
    a[J][J] += t * apq;
    a[I][J] = a[J][I] = 0.0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        if (r == I || r == J) continue;
        const double arp = a[r][I], arq = a[r][J];
        a[r][I] = a[I][r] = c * arp - sn * arq;
        a[r][J] = a[J][r] = sn * arp + c * arq;
    }
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int m = 1; m < GTA_WAVE; m <<= 1) v += __shfl_xor(v, m, GTA_WAVE);
    return v;
}

__global__ __launch_bounds__(GTA_WAVE) void k_rmsd_fp64(DevChunk Q, DevChunk R, DevState S) {
    const int p = blockIdx.x;
    if (p >= S.nq * S.nr || (S.flags[p] & CONV_LOWTMSC)) return;
    const int L = S.nalnposs[p];
    if (L < 1) return;
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    const size_t pl = S.alnplane;
    const float* g = S.aln + (size_t)qi * (size_t)R.nposs + (size_t)R.dst[ri];
    double s[17];
#pragma unroll
    for (int i = 0; i < 17; i++) s[i] = 0.0;
    for (int k = lane_id(); k < L; k += GTA_WAVE) {
        const double x = g[k], y = g[pl + k], z = g[2 * pl + k];
        const double u = g[3 * pl + k], v = g[4 * pl + k], w = g[5 * pl + k];
        s[0] += x; s[1] += y; s[2] += z; s[3] += u; s[4] += v; s[5] += w;
        s[6] += x * u; s[7] += x * v; s[8] += x * w;
        s[9] += y * u; s[10] += y * v; s[11] += y * w;
        s[12] += z * u; s[13] += z * v; s[14] += z * w;
        s[15] += x * x + y * y + z * z;
        s[16] += u * u + v * v + w * w;
    }
#pragma unroll
    for (int i = 0; i < 17; i++) s[i] = wave_sum_d(s[i]);
    const double n = (double)L;
    // centred covariance C[a][b] = Σ (q_a - q̄_a)(r_b - r̄_b) and the sums of squares
    double C[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) C[i][j] = s[6 + 3 * i + j] - s[i] * s[3 + j] / n;
    const double sq = s[15] - (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) / n;
    const double sr = s[16] - (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]) / n;
    double a[4][4];
    a[0][0] = C[0][0] + C[1][1] + C[2][2];
    a[1][1] = C[0][0] - C[1][1] - C[2][2];
    a[2][2] = -C[0][0] + C[1][1] - C[2][2];
    a[3][3] = -C[0][0] - C[1][1] + C[2][2];
    a[0][1] = a[1][0] = C[1][2] - C[2][1];
    a[0][2] = a[2][0] = C[2][0] - C[0][2];
    a[0][3] = a[3][0] = C[0][1] - C[1][0];
    a[1][2] = a[2][1] = C[0][1] + C[1][0];
    a[1][3] = a[3][1] = C[2][0] + C[0][2];
    a[2][3] = a[3][2] = C[1][2] + C[2][1];
    const double scale = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]) + fabs(a[3][3]) + 1e-300;
    for (int sweep = 0; sweep < 32; sweep++) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[0][3]) + fabs(a[1][2]) + fabs(a[1][3]) +
                           fabs(a[2][3]);
        if (off <= 1e-18 * scale) break;
        jacobi_rot<0, 1>(a); jacobi_rot<0, 2>(a); jacobi_rot<0, 3>(a);
        jacobi_rot<1, 2>(a); jacobi_rot<1, 3>(a); jacobi_rot<2, 3>(a);
    }
    const double lmax = fmax(fmax(a[0][0], a[1][1]), fmax(a[2][2], a[3][3]));
    const double e = (sq + sr - 2.0 * lmax) / n;
    if (lane_id() == 0) S.out[(size_t)p * 15 + 8] = (float)sqrt(e > 0.0 ? e : 0.0);
}

// ============================================================================
__global__ void k_setlow(DevChunk Q, DevChunk R, DevState S, float thr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.nq * S.nr) return;
    const int qi = pair_q(S, p), ri = pair_r(S, p);
    if (S.grand[p] < thr * (float)min(Q.len[qi], R.len[ri])) S.flags[p] |= CONV_LOWTMSC;
}

__global__ void k_revert(DevState S) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.nq * S.nr) return;
    float* t = S.tfm_grand + (size_t)p * 12;
    float u[12];
    copy12(u, t);
    float r[9] = {u[0], u[3], u[6], u[1], u[4], u[7], u[2], u[5], u[8]};
    for (int i = 0; i < 9; i++) t[i] = r[i];
    t[9] = -(r[0] * u[9] + r[1] * u[10] + r[2] * u[11]);
    t[10] = -(r[3] * u[9] + r[4] * u[10] + r[5] * u[11]);
    t[11] = -(r[6] * u[9] + r[7] * u[10] + r[8] * u[11]);
}

}  // namespace gta

// ============================================================================
// launchers (host side), called by the C ABI implementation
#include "gta_launch.h"

namespace gta {

static inline int blocks_for(long long n, int t) { return (int)((n + t - 1) / t); }

hipError_t launch_secstr(const DevChunk& C, hipStream_t st) {
    hipLaunchKernelGGL(k_secstr, dim3(blocks_for(C.nposs, 256)), dim3(256), 0, st, C);
    hipLaunchKernelGGL(k_secstr_na, dim3(blocks_for(C.nposs, 256)), dim3(256), 0, st, C);
    return hipGetLastError();
}
// pair_score over caller-given operands (tests of the division)
__global__ void k_pair_score(const float* d02, const float* d2, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pair_score(d02[i], d2[i]);
}
hipError_t launch_pair_score(const float* d02, const float* d2, int n, float* out, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_pair_score, dim3(blocks_for(n, 256)), dim3(256), 0, st, d02, d2, n, out);
    return hipGetLastError();
}
// VALU issue-rate calibration for the instruction counters: the first `lanes`
// lanes of every wave run `iters` iterations of 8 independent chains of one
// instruction kind K, so each wave issues exactly 8 VALU instructions of that
// kind per iteration (loop control is scalar) plus a fixed prologue and
// epilogue.  K = 0: v_fma_f32 as the compiler forms it (the factor in a scalar
// register); 1-7 one instruction each, written in inline assembly with
// register operands only: 1 v_fma_f32, 2 v_mul_f32, 3 v_add_u32, 4 v_max_f32,
// 5 v_add_f32 with a DPP row shift, 6 v_exp_f32, 7 v_mov_b32.  A grid of many
// such waves keeps every SIMD's VALU busy; the measured rates are the issue
// ceilings per instruction kind that the stage kernels are compared against.
#define GTA_CAL8(S) S(x0) S(x1) S(x2) S(x3) S(x4) S(x5) S(x6) S(x7)
#define GTA_CAL_FMA(x) asm volatile("v_fma_f32 %0, %0, %1, 1.0" : "+v"(x) : "v"(a));
#define GTA_CAL_MUL(x) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(x) : "v"(a));
#define GTA_CAL_ADDU(x) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x) : "v"(a));
#define GTA_CAL_MAX(x) asm volatile("v_max_f32 %0, %0, %1" : "+v"(x) : "v"(a));
#define GTA_CAL_DPP(x) asm volatile("v_add_f32_dpp %0, %0, %0 row_shr:1 bound_ctrl:0" : "+v"(x));
#define GTA_CAL_EXP(x) asm volatile("v_exp_f32 %0, %0" : "+v"(x));
#define GTA_CAL_MOV(x) asm volatile("v_mov_b32 %0, %1" : "=v"(x) : "v"(x));
template <int K>
__global__ __launch_bounds__(256) void k_valu_calib(float* out, int iters, float a, int lanes) {
    float x0 = (float)threadIdx.x, x1 = x0 + 1.f, x2 = x0 + 2.f, x3 = x0 + 3.f;
    float x4 = x0 + 4.f, x5 = x0 + 5.f, x6 = x0 + 6.f, x7 = x0 + 7.f;
    if ((int)(threadIdx.x & (GTA_WAVE - 1)) < lanes) {
        for (int i = 0; i < iters; i++) {
            if (K == 0) {
                x0 = __builtin_fmaf(x0, a, 1.f); x1 = __builtin_fmaf(x1, a, 1.f);
                x2 = __builtin_fmaf(x2, a, 1.f); x3 = __builtin_fmaf(x3, a, 1.f);
                x4 = __builtin_fmaf(x4, a, 1.f); x5 = __builtin_fmaf(x5, a, 1.f);
                x6 = __builtin_fmaf(x6, a, 1.f); x7 = __builtin_fmaf(x7, a, 1.f);
            }
            if (K == 1) { GTA_CAL8(GTA_CAL_FMA) }
            if (K == 2) { GTA_CAL8(GTA_CAL_MUL) }
            if (K == 3) { GTA_CAL8(GTA_CAL_ADDU) }
            if (K == 4) { GTA_CAL8(GTA_CAL_MAX) }
            if (K == 5) { GTA_CAL8(GTA_CAL_DPP) }
            if (K == 6) { GTA_CAL8(GTA_CAL_EXP) }
            if (K == 7) { GTA_CAL8(GTA_CAL_MOV) }
        }
    }
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7));
}
#undef GTA_CAL8
#undef GTA_CAL_FMA
#undef GTA_CAL_MUL
#undef GTA_CAL_ADDU
#undef GTA_CAL_MAX
#undef GTA_CAL_DPP
#undef GTA_CAL_EXP
#undef GTA_CAL_MOV
hipError_t launch_valu_calib(float* out, int blocks, int iters, int kind, int lanes, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
#define GTA_KC(K) hipLaunchKernelGGL(k_valu_calib<K>, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f, lanes)
    switch (kind) {
        case 0: GTA_KC(0); break;
        case 1: GTA_KC(1); break;
        case 2: GTA_KC(2); break;
        case 3: GTA_KC(3); break;
        case 4: GTA_KC(4); break;
        case 5: GTA_KC(5); break;
        case 6: GTA_KC(6); break;
        case 7: GTA_KC(7); break;
        default: return hipErrorInvalidValue;
    }
#undef GTA_KC
    return hipGetLastError();
}
hipError_t launch_keep_refs(const DevChunk& Q, const DevChunk& R, const DevState& S, int qgroup,
                            hipStream_t st) {
    const int g = (qgroup > 0 && qgroup < S.nq) ? qgroup : S.nq;
    const long long n = (long long)((S.nq + g - 1) / g) * S.nr;
    if (n > 0) hipLaunchKernelGGL(k_keep_refs, dim3(blocks_for(n, 256)), dim3(256), 0, st, Q, R, S, g);
    return hipGetLastError();
}
hipError_t launch_init(const DevChunk& Q, const DevChunk& R, const DevState& S, hipStream_t st) {
    const long long np = (long long)S.nq * S.nr;
    hipLaunchKernelGGL(k_init, dim3(blocks_for(np, 256)), dim3(256), 0, st, Q, R, S);
    return hipGetLastError();
}
// Part records of split pairs merged (highest score, lowest slot on ties) and
// the kernel's epilogue applied: KIND 0 seed search, 1 gapless refinement,
// 2 DP refinement in mode MODE.  One thread per split pair.
template <int KIND, int MODE>
__global__ void k_merge_parts(DevChunk Q, DevChunk R, DevState S, RefineParams prm, const int* plist,
                              int nitems, int nparts, const SlotBest* parts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    const int p = plist[i];
    const int flags = S.flags[p];
    if (KIND < 2 && (flags & CONV_LOWTMSC)) {
        S.best0[p] = 0.f;
        return;
    }
    if (KIND == 2 && MODE != RM_DPREFINE && (flags & CONV_LOWTMSC)) return;
    SlotBest b = parts[(size_t)i * nparts];
    for (int k = 1; k < nparts; k++) {
        const SlotBest& c = parts[(size_t)i * nparts + k];
        if (b.beats(c.best, c.slot)) b = c;
    }
    if (KIND == 0) findfrag_epilogue(S, p, b);
    else if (KIND == 1) gapless_epilogue(S, p, b);
    else refine_dp_epilogue<MODE>(Q, R, S, prm, p, b);
}

static int nparts_for(int nslots, int per) { return max(1, (nslots + per - 1) / per); }
size_t part_record_bytes() { return sizeof(SlotBest); }

hipError_t launch_findfrag(const DevChunk& Q, const DevChunk& R, const DevState& S, const Geom& G,
                           int dinv, const RefSplit& Sp, hipStream_t st) {
    const int nsmall = Sp.small ? Sp.nsmall : S.nq * S.nr;
    if (nsmall > 0) {
        hipLaunchKernelGGL(k_findfrag<0>, dim3(nsmall), dim3(FTHREADS), 0, st, Q, R, S, G, dinv, Sp.small, nullptr);
    }
    if (Sp.nbig > 0) {
        const int np = nparts_for(G.nseeds, Sp.seeds_per_part);
        SlotBest* parts = (SlotBest*)Sp.partbuf;
        hipLaunchKernelGGL(k_findfrag<0>, dim3(Sp.nbig, np), dim3(FTHREADS), 0, st, Q, R, S, G, dinv, Sp.big, parts);
        RefineParams prm{};
        hipLaunchKernelGGL((k_merge_parts<0, 0>), dim3((Sp.nbig + 63) / 64), dim3(64), 0, st, Q, R, S, prm, Sp.big,
                           Sp.nbig, np, parts);
    }
    return hipGetLastError();
}
hipError_t launch_refine_gapless(const DevChunk& Q, const DevChunk& R, const DevState& S,
                                 const Geom& G, int dinv, const RefSplit& Sp, hipStream_t st) {
    const int nsmall = Sp.small ? Sp.nsmall : S.nq * S.nr;
    if (nsmall > 0) {
        hipLaunchKernelGGL(k_refine_gapless<0>, dim3(nsmall), dim3(RTHREADS), 0, st, Q, R, S, G, dinv, Sp.small, nullptr);
    }
    if (Sp.nbig > 0) {
        const int np = nparts_for(G.nlocsteps, Sp.slots_per_part);
        SlotBest* parts = (SlotBest*)Sp.partbuf;
        hipLaunchKernelGGL(k_refine_gapless<0>, dim3(Sp.nbig, np), dim3(RTHREADS), 0, st, Q, R, S, G, dinv, Sp.big, parts);
        RefineParams prm{};
        hipLaunchKernelGGL((k_merge_parts<1, 0>), dim3((Sp.nbig + 63) / 64), dim3(64), 0, st, Q, R, S, prm, Sp.big,
                           Sp.nbig, np, parts);
    }
    return hipGetLastError();
}
namespace {
struct LengthTables {
    std::vector<float> cb, pw;
    LengthTables() : cb(MAXLEN + 1), pw(MAXLEN + 1) {
        for (int L = 0; L <= MAXLEN; L++) {
            volatile float l = (float)L;      // evaluated at run time by the C library
            cb[L] = cbrtf(l - 15.f);
            pw[L] = powf(l, 0.3f);
        }
    }
};
}  // namespace

// the residue score table in the device layout (codes 0..25 for 'A'..'Z', 26
// for any other character)
struct ResidueTable {
    float g[NGT * NGT];
    ResidueTable() {
        float bits[GONNET_NA * GONNET_NA];
        gonnet_bits(bits);
        for (int a = 0; a < NGT; a++)
            for (int b = 0; b < NGT; b++)
                g[a * NGT + b] = (a < GONNET_NA && b < GONNET_NA) ? bits[a * GONNET_NA + b] : 0.0f;
    }
};

hipError_t upload_length_tables() {
    static const LengthTables t;              // built once (thread-safe static)
    static const ResidueTable rt;
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g_cbrt_lm15), t.cb.data(), t.cb.size() * sizeof(float));
    if (e != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(g_gonnet), rt.g, sizeof rt.g)) != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(g_pow03), t.pw.data(), t.pw.size() * sizeof(float));
}

hipError_t launch_dp(const DevChunk& Q, const DevChunk& R, const DevState& S, const DpLaunch& L,
                     const DpBuckets& B, const WorkQ& wq, int pgrid, hipStream_t st) {
    DpParams prm{L.mode, L.gap0, L.gapopen, L.usess, L.ssw, L.dpscan, L.rrw};
    // kernels specialised on (score kind, gap mode); the SS/residue fill
    // always has a gap cost
#define GTA_DPV(KCALL)                                                      \
    do {                                                                    \
        if (L.usess == SC_SSRR) {                                           \
            KCALL(SC_SSRR, false);                                          \
        } else if (L.usess == SC_TMSS) {                                    \
            if (L.gap0) KCALL(SC_TMSS, true); else KCALL(SC_TMSS, false);   \
        } else {                                                            \
            if (L.gap0) KCALL(SC_TM, true); else KCALL(SC_TM, false);       \
        }                                                                   \
    } while (0)
    const int nsmall = wq.items ? pgrid : (B.small ? B.nsmall : S.nq * S.nr);
    if (nsmall > 0) {
#define GTA_K1(US, G0) \
    hipLaunchKernelGGL((k_dp<US, G0>), dim3(nsmall), dim3(GTA_WAVE), 0, st, Q, R, S, prm, B.small, wq)
        GTA_DPV(GTA_K1);
#undef GTA_K1
    }
#define GTA_MW(US, G0, WW, G)                                                                                \
    hipLaunchKernelGGL((k_dp_mw<US, G0, WW>), dim3(B.nmw[G]), dim3(WW * GTA_WAVE), 0, st, Q, R, S, prm, B.mw[G], \
                       B.ring, B.mwoff[G], B.failflag)
#define GTA_KMW4(US, G0) GTA_MW(US, G0, 4, 0)
#define GTA_KMW8(US, G0) GTA_MW(US, G0, 8, 1)
#define GTA_KMW16(US, G0) GTA_MW(US, G0, 16, 2)
    if (B.nmw[0] > 0) GTA_DPV(GTA_KMW4);
    if (B.nmw[1] > 0) GTA_DPV(GTA_KMW8);
    if (B.nmw[2] > 0) GTA_DPV(GTA_KMW16);
#undef GTA_KMW4
#undef GTA_KMW8
#undef GTA_KMW16
#undef GTA_MW
    if (B.ngrid > 0 && L.usess == SC_SSRR) {
        // The SS/residue fill does not use the device-wide pipeline: its
        // grid-form run stalled on the GPU with no cause found yet.  Those
        // pairs run as 4-wave workgroups instead, on their grid rows (a pair
        // holds nstrips >= min(4, nstrips) rows of Lr floats there).
        hipLaunchKernelGGL((k_dp_mw<SC_SSRR, false, 4>), dim3(B.ngrid), dim3(4 * GTA_WAVE), 0, st, Q, R, S, prm,
                           B.grid, B.ring, B.growoff, B.failflag);
    } else if (B.ngrid > 0) {
        // progress counters and the item counter start at zero for every fill
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for((long long)B.gtotal + 1, 256)), dim3(256), 0, st, B.gsync, 0,
                           (size_t)B.gtotal + 1);
#define GTA_KG(US, G0)                                                                                    \
    hipLaunchKernelGGL((k_dp_grid<US, G0>), dim3(B.gworkers), dim3(GTA_WAVE), 0, st, Q, R, S, prm, B.grid, \
                       B.ngrid, B.gstrip0, B.gsync, B.ring, B.growoff, B.failflag)
        GTA_DPV(GTA_KG);
#undef GTA_KG
        hipLaunchKernelGGL(k_dp_grid_walk, dim3(B.ngrid), dim3(GTA_WAVE), 0, st, Q, R, S, prm, B.grid, B.failflag);
    }
#undef GTA_DPV
    return hipGetLastError();
}
hipError_t launch_dp_score(const DevChunk& Q, const DevChunk& R, const DevState& S, int gap0, float gapopen,
                           const float* tfms, const int* slotflags, int nslots, float* scores, hipStream_t st) {
    const long long np = (long long)S.nq * S.nr;
    if (np > 0 && nslots > 0) {
        if (gap0)
            hipLaunchKernelGGL(k_dp_score<true>, dim3((unsigned)np), dim3(GTA_WAVE), 0, st, Q, R, S, gapopen, tfms,
                               slotflags, nslots, scores);
        else
            hipLaunchKernelGGL(k_dp_score<false>, dim3((unsigned)np), dim3(GTA_WAVE), 0, st, Q, R, S, gapopen, tfms,
                               slotflags, nslots, scores);
    }
    return hipGetLastError();
}
hipError_t launch_refine_dp(int prec, int mode, const DevChunk& Q, const DevChunk& R, const DevState& S,
                            const Geom& G, const RefineLaunch& L, const RefSplit& Sp, const WorkQ& wq,
                            int pgrid, hipStream_t st) {
    RefineParams prm{L.dinv, L.epi, L.dpi, L.maxndpiters, L.prescore};
    // DPRefine runs in workgroups of L.waves (1 or 2), the other modes in 2-wave ones
    const bool w1 = mode == RM_DPREFINE && L.waves == 1;
    const int ldscap = (prec == 0 || mode == RM_DPREFINE) ? std::min(L.ldscap, rlds_cap32(w1 ? 1 : RWAVES))
                                                          : L.ldscap;
    const size_t lds = (size_t)ldscap * 6 * sizeof(float);
    const int nsmall = wq.items ? pgrid : (Sp.small ? Sp.nsmall : S.nq * S.nr);
    const bool phase2 = (mode == RM_FIN2 || mode == RM_PROD2);
    const bool full2 = (mode == RM_FIN2F || mode == RM_PROD2F);
    const int np = Sp.nbig > 0 ? nparts_for(phase2 ? G.maxnfragfcts
                                                   : full2 ? G.maxnfragfcts * NMAXSUBFRAGS : G.nlocsteps,
                                            Sp.slots_per_part)
                               : 1;
    SlotBest* parts = (SlotBest*)Sp.partbuf;
#define GTA_RL(PP, MM)                                                                                        \
    do {                                                                                                      \
        if (nsmall > 0)                                                                                       \
            hipLaunchKernelGGL((k_refine_dp<PP, MM>), dim3(nsmall), dim3(RTHREADS), lds, st, Q, R, S, G, prm, \
                               ldscap, Sp.small, (SlotBest*)nullptr, wq);                                   \
        if (Sp.nbig > 0) {                                                                                    \
            hipLaunchKernelGGL((k_refine_dp<PP, MM>), dim3(Sp.nbig, np), dim3(RTHREADS), lds, st, Q, R, S, G, \
                               prm, ldscap, Sp.big, parts, WorkQ{});                                        \
            hipLaunchKernelGGL((k_merge_parts<2, MM>), dim3((Sp.nbig + 63) / 64), dim3(64), 0, st, Q, R, S,  \
                               prm, Sp.big, Sp.nbig, np, parts);                                              \
        }                                                                                                     \
    } while (0)
    // the same with 1-wave workgroups
#define GTA_RL1(PP, MM)                                                                                        \
    do {                                                                                                       \
        if (nsmall > 0)                                                                                        \
            hipLaunchKernelGGL((k_refine_dp<PP, MM, 1>), dim3(nsmall), dim3(GTA_WAVE), lds, st, Q, R, S, G, prm, \
                               ldscap, Sp.small, (SlotBest*)nullptr, wq);                                    \
        if (Sp.nbig > 0) {                                                                                     \
            hipLaunchKernelGGL((k_refine_dp<PP, MM, 1>), dim3(Sp.nbig, np), dim3(GTA_WAVE), lds, st, Q, R, S, G, \
                               prm, ldscap, Sp.big, parts, WorkQ{});                                         \
            hipLaunchKernelGGL((k_merge_parts<2, MM>), dim3((Sp.nbig + 63) / 64), dim3(64), 0, st, Q, R, S,   \
                               prm, Sp.big, Sp.nbig, np, parts);                                               \
        }                                                                                                      \
    } while (0)
    // the log search's steps run in order within one workgroup: no split
#define GTA_RL_LOG(PP)                                                                                         \
    do {                                                                                                       \
        if (nsmall > 0)                                                                                        \
            hipLaunchKernelGGL((k_refine_dp<PP, RM_PROD2L>), dim3(nsmall), dim3(RTHREADS), lds, st, Q, R, S, G, \
                               prm, ldscap, Sp.small, (SlotBest*)nullptr, wq);                               \
        if (Sp.nbig > 0)                                                                                       \
            hipLaunchKernelGGL((k_refine_dp<PP, RM_PROD2L>), dim3(Sp.nbig), dim3(RTHREADS), lds, st, Q, R, S, G,\
                               prm, ldscap, Sp.big, (SlotBest*)nullptr, WorkQ{});                            \
    } while (0)
#define GTA_RLM(PP)                                      \
    switch (mode) {                                      \
        case RM_DPREFINE:                                \
            if (w1) GTA_RL1(0, RM_DPREFINE);             \
            else GTA_RL(0, RM_DPREFINE);                 \
            break;                                       \
        case RM_FIN1: GTA_RL(PP, RM_FIN1); break;         \
        case RM_FIN2: GTA_RL(PP, RM_FIN2); break;         \
        case RM_FIN2F: GTA_RL(PP, RM_FIN2F); break;       \
        case RM_PROD1: GTA_RL(PP, RM_PROD1); break;       \
        case RM_PROD2F: GTA_RL(PP, RM_PROD2F); break;     \
        case RM_PROD2L: GTA_RL_LOG(PP); break;            \
        default: GTA_RL(PP, RM_PROD2); break;             \
    }
    if (prec) { GTA_RLM(1) } else { GTA_RLM(0) }
#undef GTA_RLM
#undef GTA_RL_LOG
#undef GTA_RL1
#undef GTA_RL
    return hipGetLastError();
}
hipError_t launch_prod_aln(const DevChunk& Q, const DevChunk& R, const DevState& S, char* alns, int nodel,
                           const WorkQ& wq, int pgrid, hipStream_t st) {
    const int n = wq.items ? pgrid : S.nq * S.nr;
    hipLaunchKernelGGL(k_prod_aln, dim3(n), dim3(GTA_WAVE), 0, st, Q, R, S, alns, nodel, wq);
    return hipGetLastError();
}
hipError_t launch_live_lists(const DevState& S, const unsigned char* bigmask, int* cnt, int* totals, int* lista,
                             int* listb, int* listc, hipStream_t st) {
    const long long np = (long long)S.nq * S.nr;
    const int nblk = (int)((np + LBLK - 1) / LBLK);
    hipLaunchKernelGGL(k_live_count, dim3(nblk), dim3(LBLK), 0, st, S, bigmask, cnt);
    hipLaunchKernelGGL(k_live_scan, dim3(1), dim3(64), 0, st, cnt, nblk, totals);
    hipLaunchKernelGGL(k_live_scatter, dim3(nblk), dim3(LBLK), 0, st, S, bigmask, cnt, lista, listb, listc);
    return hipGetLastError();
}
hipError_t launch_fill_i32(int* d, int v, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for((long long)n, 256)), dim3(256), 0, st, d, v, n);
    return hipGetLastError();
}
hipError_t launch_rmsd_fp64(const DevChunk& Q, const DevChunk& R, const DevState& S, hipStream_t st) {
    const long long np = (long long)S.nq * S.nr;
    if (np > 0) hipLaunchKernelGGL(k_rmsd_fp64, dim3((unsigned)np), dim3(GTA_WAVE), 0, st, Q, R, S);
    return hipGetLastError();
}
hipError_t launch_setlow(const DevChunk& Q, const DevChunk& R, const DevState& S, float thr, hipStream_t st) {
    const long long np = (long long)S.nq * S.nr;
    hipLaunchKernelGGL(k_setlow, dim3(blocks_for(np, 256)), dim3(256), 0, st, Q, R, S, thr);
    return hipGetLastError();
}
hipError_t launch_revert(const DevState& S, hipStream_t st) {
    const long long np = (long long)S.nq * S.nr;
    hipLaunchKernelGGL(k_revert, dim3(blocks_for(np, 256)), dim3(256), 0, st, S);
    return hipGetLastError();
}

}  // namespace gta
