// gta_launch.h -- host-side launchers of the stage kernels (gta_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include "gta_device.h"

namespace gta {

enum { DPM_IBEST_L = 0, DPM_GRAND_L = 1, DPM_CONSTRAINED_L = 2 };
enum { RM_DPREFINE_L = 0, RM_FIN1_L = 1, RM_FIN2_L = 2, RM_PROD1_L = 3, RM_PROD2_L = 4 };

struct DpLaunch {
    int mode;          // DPM_*_L
    int gap0;
    float gapopen;
    int usess;
    float ssw;
    int dpscan;
};

struct RefineLaunch {
    int dinv;
    int epi, dpi, maxndpiters;
    float prescore;
    int ldscap;        // alignments up to this length are staged in LDS
};

hipError_t launch_secstr(const DevChunk& C, hipStream_t st);
hipError_t launch_init(const DevChunk& Q, const DevChunk& R, const DevState& S, hipStream_t st);
hipError_t launch_findfrag(int prec, const DevChunk& Q, const DevChunk& R, const DevState& S,
                           const Geom& G, int dinv, hipStream_t st);
hipError_t launch_refine_gapless(int prec, const DevChunk& Q, const DevChunk& R, const DevState& S,
                                 const Geom& G, int dinv, hipStream_t st);
// per-length d0/d8 tables on the current device (once per device, before any launch)
hipError_t upload_length_tables();
hipError_t launch_dp(const DevChunk& Q, const DevChunk& R, const DevState& S, const DpLaunch& L,
                     hipStream_t st);
hipError_t launch_refine_dp(int prec, int mode, const DevChunk& Q, const DevChunk& R,
                            const DevState& S, const Geom& G, const RefineLaunch& L, hipStream_t st);
hipError_t launch_prod_aln(const DevChunk& Q, const DevChunk& R, const DevState& S, char* alns,
                           int nodel, hipStream_t st);
hipError_t launch_setlow(const DevChunk& Q, const DevChunk& R, const DevState& S, float thr,
                         hipStream_t st);
hipError_t launch_revert(const DevState& S, hipStream_t st);

}  // namespace gta
