/*
 * gtalign_amd_testing.h -- test hooks of the stage library.
 *
 * No counterpart in the reference: these entry points exist so that tests
 * can drive single kernels of the stage on given inputs (the reference's
 * equivalent checks are compile-time TESTPRINT dumps inside its kernels).
 */
#ifndef GTALIGN_AMD_TESTING_H
#define GTALIGN_AMD_TESTING_H

#include "gtalign_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* DP fill + backtrace (the kernel behind ExecDPTFMSSwBtck128xKernel and
 * BtckToMatched128xKernel, MpDPHub.cpp:687-984,1500-1610) for every pair of
 * the resident chunks under the caller's per-pair transforms tfms[pair*12].
 * d0 is the search d0 (dpscan=0) or the DP-scan d0 (dpscan=1); usess=1 adds
 * 0.5 for equal secondary-structure letters.  usess=2 runs instead the fill
 * of the --add-search-by-ss search (ExecDPSSwBtck128xKernel<true>,
 * MpDPHub.cpp:985-1230, as MpStageSSRR::Align calls it): score 1 for equal
 * SS letters plus 0.2 x the residue score, gap open gapopen, no transform
 * used, gap0 ignored.  nmatched[pair] receives the
 * number of matched pairs; qpos/rpos receive the matched positions in walk
 * order at offset q*nrefposs + dst[r] for pair (q, r) (arrays of
 * nqrys*nrefposs ints). */
int gta_test_dp(gta_ctx* ctx, const float* tfms, int gap0, float gapopen, int usess, int dpscan,
                int* nmatched, int* qpos, int* rpos);

/* Score-only DP fills (the kernel behind ExecDPScore3264x / the CPU twin's
 * ExecDPScore128xKernel, MpDPHub.cpp:389-647, which the frg3 superposition
 * search uses to rank candidate transforms) of nslots superpositions per pair:
 * slot k of pair p = qi*nrefs + ri uses tfms[(k*npairs + p)*12] and
 * scores[k*npairs + p] receives the value of the fill's bottom-right cell
 * (the score the reference reads back, MpStageFrg3.h:1186-1191), or 0 for a
 * slot whose slotflags entry (may be NULL) is nonzero.  Search d0. */
int gta_test_dp_score(gta_ctx* ctx, const float* tfms, const int* slotflags, int nslots, int gap0,
                      float gapopen, float* scores);

/* Thresholds of the DP's size buckets: a pair is filled by a workgroup of
 * several waves (strips pipelined down the query) when Lq*Lr >= min_cells,
 * it has at least min_strips 64-row strips and Lr >= 256; otherwise by one
 * wave.  Defaults 4,194,304 cells and 16 strips.  Takes effect at the next
 * gta_process / gta_test_dp.  Lets tests drive the multi-wave fill on
 * moderate sizes. */
int gta_test_set_dp_buckets(gta_ctx* ctx, long long min_cells, int min_strips);

/* Pairs of the multi-wave bucket above are filled by the whole device
 * instead (one-wave workgroups pipelining the pair's strips through global
 * memory) when mode is 1, never when 2, and when they number at most half
 * the device's compute units when 0 (the default).  Takes effect at the next
 * gta_process / gta_test_dp. */
int gta_test_set_dp_grid(gta_ctx* ctx, int mode);

/* DP launch forms of the resident chunk pair as of the last gta_process or
 * gta_test_dp: counts[0..4] receive the number of pairs filled by one wave,
 * by 4, 8 and 16 waves and by the device-wide pipeline; tune_seconds (may be
 * NULL) the host time spent autotuning so far. */
int gta_test_dp_forms(const gta_ctx* ctx, int* counts, double* tune_seconds);

/* Split of the superposition searches over workgroups: a pair whose shorter
 * structure has at least min_len residues has its seed diagonals
 * (seeds_per_part per workgroup) and fragment slots (slots_per_part per
 * workgroup) spread over several workgroups, whose best slots are merged by
 * a second kernel.  Defaults 2048, 256 and 32.  Takes effect at the next
 * gta_process. */
int gta_test_set_split(gta_ctx* ctx, int min_len, int seeds_per_part, int slots_per_part);

/* The pair score d0^2 / (d0^2 + d^2) of the DP cells and TM-score passes
 * (scoringbase.h:61-65) for n caller-given operand pairs: out[i] receives
 * the device's value for d02[i], d2[i] (its division must round as IEEE
 * division does). */
int gta_test_pair_score(gta_ctx* ctx, const float* d02, const float* d2, int n, float* out);

/* VALU issue-rate calibration for the instruction counters: `blocks`
 * workgroups of 256 lanes; the first `lanes` (1-64) lanes of every wave run
 * `iters` iterations of 8 independent chains of one instruction kind:
 * 0 v_fma_f32 (compiler-formed, factor in a scalar register), 1 v_fma_f32,
 * 2 v_mul_f32, 3 v_add_u32, 4 v_max_f32, 5 v_add_f32 with a DPP row shift,
 * 6 v_exp_f32, 7 v_mov_b32 (1-7 with vector register operands).  Every wave
 * issues exactly 8 * iters VALU instructions of the kind in the loop plus 25
 * (kinds 0, 5, 6, 7) or 26 (kinds 1-4) before and after it.  `seconds`
 * receives the launch's duration (HIP events on the context's stream, after
 * one untimed launch). */
int gta_test_valu_calibration(gta_ctx* ctx, int blocks, int iters, int kind, int lanes, double* seconds);

#ifdef __cplusplus
}
#endif

#endif /* GTALIGN_AMD_TESTING_H */
