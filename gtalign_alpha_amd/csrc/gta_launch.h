// gta_launch.h -- host-side launchers of the stage kernels (gta_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include "gta_device.h"

namespace gta {

enum { DPM_IBEST_L = 0, DPM_GRAND_L = 1, DPM_CONSTRAINED_L = 2 };
enum { RM_DPREFINE_L = 0, RM_FIN1_L = 1, RM_FIN2_L = 2, RM_PROD1_L = 3, RM_PROD2_L = 4, RM_FIN2F_L = 5,
       RM_PROD2F_L = 6, RM_PROD2L_L = 7 };

struct DpLaunch {
    int mode;          // DPM_*_L
    int gap0;
    float gapopen;
    int usess;         // cell score: 0 distance, 1 distance + SS bonus, 2 SS match + residue similarity
    float ssw;
    int dpscan;
    float rrw;         // residue-score weight of kind 2
};

// Pairs of a chunk split by DP size: the one-wave kernel takes `small`
// (nullptr: every pair), the W-wave kernel takes `big` with W x Lr floats
// of ring rows per pair at ringoff[i].
struct DpBuckets {
    const int* small;                       // one wave per pair (nullptr: every pair)
    int nsmall;
    // one workgroup of W = 4 << g waves per pair, g = 0, 1, 2; ring rows of
    // W x Lr floats per pair at mwoff[g][k] in ring
    const int* mw[3];
    int nmw[3];
    const unsigned long long* mwoff[3];
    float* ring;
    int* failflag;     // set by a W-wave or grid fill whose row wait gave up
    // pairs filled by the device-wide strip pipeline (k_dp_grid)
    const int* grid;
    int ngrid;
    const int* gstrip0;                     // [ngrid + 1] first strip item of each pair
    int gtotal;                             // strips of all grid pairs
    int* gsync;                             // [gtotal + 1] strip progress, then the item counter
    const unsigned long long* growoff;      // per pair nstrips x Lr floats at growoff[k] in ring
    int gworkers;                           // one-wave workgroups of the fill
    // any pair outside the one-wave list
    bool any_big() const { return nmw[0] > 0 || nmw[1] > 0 || nmw[2] > 0 || ngrid > 0; }
};

// Pairs of a chunk split by refinement size: `small` pairs (nullptr: every
// pair) take one workgroup each; each `big` pair's slots are split over
// several workgroups whose part records (partbuf, nbig x parts entries) are
// merged by a second kernel.
struct RefSplit {
    const int* small;
    int nsmall;
    const int* big;
    int nbig;
    int seeds_per_part;    // seed diagonals per workgroup in the seed search
    int slots_per_part;    // fragment slots per workgroup in the refinements
    void* partbuf;
};

struct RefineLaunch {
    int dinv;
    int epi, dpi, maxndpiters;
    float prescore;
    int ldscap;        // alignments up to this length are staged in LDS
    int waves;         // waves per workgroup of DPRefine (1 or 2; the other modes use 2)
};

hipError_t launch_secstr(const DevChunk& C, hipStream_t st);
hipError_t launch_init(const DevChunk& Q, const DevChunk& R, const DevState& S, hipStream_t st);
// the pair score d02 / (d02 + d2) over device arrays of n operands
hipError_t launch_pair_score(const float* d02, const float* d2, int n, float* out, hipStream_t st);
// the VALU calibration kernel of instruction kind `kind` (0-7) on the first
// `lanes` lanes of every wave
hipError_t launch_valu_calib(float* out, int blocks, int iters, int kind, int lanes, hipStream_t st);
// the reference filter after the stage-1 prescreen, per block of qgroup queries (0: all)
hipError_t launch_keep_refs(const DevChunk& Q, const DevChunk& R, const DevState& S, int qgroup,
                            hipStream_t st);
// the superposition searches run in the reference's FP32 arithmetic in both
// precision forms (their result is a choice of superposition, which the final
// stage refines in the precision of the form)
hipError_t launch_findfrag(const DevChunk& Q, const DevChunk& R, const DevState& S,
                           const Geom& G, int dinv, const RefSplit& Sp, hipStream_t st);
hipError_t launch_refine_gapless(const DevChunk& Q, const DevChunk& R, const DevState& S,
                                 const Geom& G, int dinv, const RefSplit& Sp, hipStream_t st);
// per-length d0/d8 tables on the current device (once per device, before any launch)
hipError_t upload_length_tables();
// wq.items != nullptr: a persistent grid of pgrid workgroups drains the
// work queue instead of the small bucket (batched form)
hipError_t launch_dp(const DevChunk& Q, const DevChunk& R, const DevState& S, const DpLaunch& L,
                     const DpBuckets& B, const WorkQ& wq, int pgrid, hipStream_t st);
// score-only DP fills of nslots superpositions per pair (scores[k*np + p])
hipError_t launch_dp_score(const DevChunk& Q, const DevChunk& R, const DevState& S, int gap0, float gapopen,
                           const float* tfms, const int* slotflags, int nslots, float* scores, hipStream_t st);
// prec selects FP64 covariance and solve for the final-stage modes (FIN1,
// FIN2, PROD1, PROD2); DPREFINE is a search and stays FP32
hipError_t launch_refine_dp(int prec, int mode, const DevChunk& Q, const DevChunk& R,
                            const DevState& S, const Geom& G, const RefineLaunch& L, const RefSplit& Sp,
                            const WorkQ& wq, int pgrid, hipStream_t st);
// bytes of one part record (partbuf entries)
size_t part_record_bytes();
hipError_t launch_prod_aln(const DevChunk& Q, const DevChunk& R, const DevState& S, char* alns,
                           int nodel, const WorkQ& wq, int pgrid, hipStream_t st);
// work lists of live pairs (batched form): lista for the one-wave DP, listb for
// one-workgroup refinement, listc (optional) every live pair for the output
// pass; cnt holds 3 ints per 1024 pairs, totals 3 ints
hipError_t launch_live_lists(const DevState& S, const unsigned char* bigmask, int* cnt, int* totals, int* lista,
                             int* listb, int* listc, hipStream_t st);
hipError_t launch_fill_i32(int* d, int v, size_t n, hipStream_t st);
// mixed precision: RMSD of every output pair's final alignment from FP64 sums
// and the quaternion matrix's largest eigenvalue (after the PROD1 refinement)
hipError_t launch_rmsd_fp64(const DevChunk& Q, const DevChunk& R, const DevState& S, hipStream_t st);
hipError_t launch_setlow(const DevChunk& Q, const DevChunk& R, const DevState& S, float thr,
                         hipStream_t st);
hipError_t launch_revert(const DevState& S, hipStream_t st);

}  // namespace gta
