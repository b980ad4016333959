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
 * d0 is the search d0 (dpscan=0) or the DP-scan d0 (dpscan=1); usess adds
 * 0.5 for equal secondary-structure letters.  nmatched[pair] receives the
 * number of matched pairs; qpos/rpos receive the matched positions in walk
 * order at offset q*nrefposs + dst[r] for pair (q, r) (arrays of
 * nqrys*nrefposs ints). */
int gta_test_dp(gta_ctx* ctx, const float* tfms, int gap0, float gapopen, int usess, int dpscan,
                int* nmatched, int* qpos, int* rpos);

#ifdef __cplusplus
}
#endif

#endif /* GTALIGN_AMD_TESTING_H */
