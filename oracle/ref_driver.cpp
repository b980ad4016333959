// ref_driver.cpp -- runs the reference's own CPU path (libmymp, compiled from
// /root/reference/src by oracle/Makefile) on a query chunk and a reference
// chunk, and returns its result sections.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle and as the timed CPU
// baseline; it is never part of the product.  The sequence below follows
// MpBatch::ProcessBlock (src/libmymp/mpbatch/MpBatch.cpp:283-485) with the
// sequence-similarity pre-filter, the frg3 search (--no-detailed-search) and
// the finalizer thread left out.  Reference compaction after stage 1
// (FilteroutReferences, MpBatch.cpp:359-367) is replaced by leaving the
// flagged pairs in the chunk: every later kernel skips pairs carrying the
// low-score flag, so the surviving pairs see the same computation.

#include <omp.h>
#include <chrono>
#include <cstring>
#include <vector>

#include "libutil/mybase.h"
#include "libutil/CLOptions.h"
#include "libgenp/gproc/gproc.h"
#include "libgenp/gdats/PM2DVectorFields.h"
#include "libgenp/gdats/PMBatchStrData.h"
#include "libmycu/custages/fragment.cuh"
#include "libmymp/mplayout/MpGlobalMemory.h"
#include "libmymp/mpss/MpSecStr.h"
#include "libmymp/mpstage1/MpStage1.h"
#include "libmymp/mpstage1/MpStage2.h"
#include "libmymp/mpstage1/MpStageFin.h"

#include "gta/gtalign_amd.h"

namespace {

struct ChunkFields {
    std::vector<char> buf[pmv2DTotFlds];
    char* beg[pmv2DTotFlds];
    char* end[pmv2DTotFlds];
};

template <typename T>
void put(std::vector<char>& b, const T* src, size_t n, T defv) {
    b.resize(n * sizeof(T) + 64);
    T* d = (T*)b.data();
    for (size_t i = 0; i < n; i++) d[i] = src ? src[i] : defv;
}

void fill(const gta_chunk_view* c, ChunkFields& f) {
    const size_t ns = c->nstrs, np = c->nposs;
    put<INTYPE>(f.buf[pps2DLen], c->len, ns, 0);
    put<LNTYPE>(f.buf[pps2DDist], c->dst, ns, 0);
    put<INTYPE>(f.buf[pps2DType], c->type, ns, 0);
    put<FPTYPE>(f.buf[pmv2DCoords + 0], c->x, np, 0.f);
    put<FPTYPE>(f.buf[pmv2DCoords + 1], c->y, np, 0.f);
    put<FPTYPE>(f.buf[pmv2DCoords + 2], c->z, np, 0.f);
    put<INTYPE>(f.buf[pmv2D_Ins_Ch_Ord], c->ins_ch_ord, np, 0);
    put<INTYPE>(f.buf[pmv2DResNumber], c->resnum, np, 0);
    put<CHTYPE>(f.buf[pmv2Drsd], c->rsd, np, 'A');
    put<CHTYPE>(f.buf[pmv2Dss], c->ss, np, ' ');
    for (int i = 0; i < pmv2DTotFlds; i++) {
        f.beg[i] = f.buf[i].data();
        size_t n = (i < pps2DStrFlds) ? ns : np;
        f.end[i] = f.beg[i] + n * TPM2DVectorFieldSize::szvfs_[i];
    }
}

}  // namespace

extern "C" {

// Options are read once per process by the reference (function-local
// statics), so a process must use one option set and one thread count.
int gta_ref_process_block(const gta_chunk_view* q, const gta_chunk_view* r, const gta_options* o,
                          int nthreads, float* alndata, float* tfms, char* alns, size_t alnsize,
                          double* seconds) {
    try {
        CLOptions::AssignCLOptCPU_THREADS(nthreads);
        CLOptions::AssignCLOptP_PRE_SCORE(o->prescore);
        CLOptions::AssignCLOptO_S(o->scorethld);
        CLOptions::AssignCLOptC_CONVERGENCE(o->convergence);
        CLOptions::AssignCLOptC_DEPTH(o->depth);
        CLOptions::AssignCLOptC_REFINEMENT(o->refinement);
        CLOptions::AssignCLOptC_SYMMETRIC(o->symmetric);
        CLOptions::AssignCLOptO_NO_DELETIONS(o->nodeletions);
        CLOptions::AssignCLOptO_REFERENCED(o->referenced);
        CLOptions::AssignCLOptC_NODETAILEDSEARCH(1);

        ChunkFields qf, rf;
        fill(q, qf);
        fill(r, rf);

        const size_t nqystrs = q->nstrs, ndbCstrs = r->nstrs;
        const size_t nqyposs = q->nposs, ndbCposs = r->nposs;
        const size_t qystr1len = q->len[0], dbstr1len = r->len[0];
        const size_t qystrnlen = q->len[nqystrs - 1], dbstrnlen = r->len[ndbCstrs - 1];
        const size_t dbxpad = ALIGN_UP(ndbCposs, CUL2CLINESIZE) - ndbCposs;

        static MpGlobalMemory* gmem = nullptr;
        static size_t gmem_qposs = 0;
        if (!gmem || gmem_qposs < nqyposs) {
            delete gmem;
            gmem = new MpGlobalMemory((size_t)8 * ONEG, 1);
            gmem->CalcMaxDbDataChunkSize(nqyposs);
            gmem_qposs = nqyposs;
        }
        if (gmem->GetCurrentMaxDbPos() < ndbCposs + dbxpad) return -4;
        auto sec = [&](int s) { return gmem->GetHeap() + gmem->GetHeapSectionOffset(0, s); };

        const uint maxnsteps = MpGlobalMemory::GetMaxNFragSteps();
        const uint minfraglen = GetMinFragLengthForAln(MpGlobalMemory::GetMaxAlnLength());
        const int stepinit = (o->depth == CLOptions::csdShallow) ? 5 : 1;

        float* scores = (float*)sec(MpGlobalMemory::ddsBegOfMtxScores);
        float* tmpdpdiagbuffers = (float*)sec(MpGlobalMemory::ddsBegOfDPDiagScores);
        float* tmpdpbotbuffer = (float*)sec(MpGlobalMemory::ddsBegOfDPBottomScores);
        float* tmpdpalnpossbuffer = (float*)sec(MpGlobalMemory::ddsBegOfDPAlignedPoss);
        uint* maxcoordsbuf = (uint*)sec(MpGlobalMemory::ddsBegOfDPMaxCoords);
        char* btckdata = sec(MpGlobalMemory::ddsBegOfDPBackTckData);
        float* wrkmemory = (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory);
        float* wrkmemoryccd = (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryCCD);
        float* wrkmemorytm = (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryTM);
        float* wrkmemorytmibest = (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryTMibest);
        float* auxwrkmemory = (float*)sec(MpGlobalMemory::ddsBegOfAuxWrkMemory);
        float* wrkmemory2 = (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory2);
        float* alndatamemory = (float*)sec(MpGlobalMemory::ddsBegOfDP2AlnData);
        float* tfmmemory = (float*)sec(MpGlobalMemory::ddsBegOfTfmMatrices);
        char* alnsmemory = sec(MpGlobalMemory::ddsBegOfDP2Alns);
        uint* globvarsbuf = (uint*)sec(MpGlobalMemory::ddsBegOfGlobVars);

        char** qb = qf.beg; char** qe = qf.end;
        char** rb = rf.beg; char** re = rf.end;
        const size_t szalns = gmem->GetSizeOfDP2Alns(nqystrs, nqyposs, ndbCposs, ndbCstrs);
        std::memset(alnsmemory, 0, szalns);

        auto t0 = std::chrono::steady_clock::now();
        MpStage1 stg1(2, maxnsteps, minfraglen, o->prescore, stepinit, qb, qe, rb, re,
                      nqystrs, ndbCstrs, nqyposs, ndbCposs, qystr1len, dbstr1len, qystrnlen, dbstrnlen,
                      dbxpad, scores, tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer, maxcoordsbuf,
                      btckdata, wrkmemory, wrkmemoryccd, wrkmemorytm, wrkmemorytmibest, auxwrkmemory,
                      wrkmemory2, alndatamemory, tfmmemory, globvarsbuf);
        stg1.Preinitialize1(false);
        MpSecStr(tmpdpdiagbuffers, qb, qe, rb, re, nqystrs, ndbCstrs, nqyposs, ndbCposs, qystr1len,
                 dbstr1len, qystrnlen, dbstrnlen, dbxpad).Run();
        stg1.Run();
        MpStage2 stg2(2, maxnsteps, minfraglen, o->prescore, stepinit, qb, qe, rb, re,
                      nqystrs, ndbCstrs, nqyposs, ndbCposs, qystr1len, dbstr1len, qystrnlen, dbstrnlen,
                      dbxpad, scores, tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer, maxcoordsbuf,
                      btckdata, wrkmemory, wrkmemoryccd, wrkmemorytm, wrkmemorytmibest, auxwrkmemory,
                      wrkmemory2, alndatamemory, tfmmemory, globvarsbuf);
        stg2.RunSpecialized<false, true, D02IND_DPSCAN>(2, false, o->scorethld);
        stg2.RunSpecialized<true, false, D02IND_SEARCH>(o->convergence, true, o->scorethld);
        MpStageFin(EQUIVALENCE_DISTANCE * EQUIVALENCE_DISTANCE, maxnsteps, minfraglen, qb, qe, rb, re,
                   nqystrs, ndbCstrs, nqyposs, ndbCposs, qystr1len, dbstr1len, qystrnlen, dbstrnlen,
                   dbxpad, scores, tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer, maxcoordsbuf,
                   btckdata, wrkmemory, wrkmemoryccd, wrkmemorytm, wrkmemorytmibest, auxwrkmemory,
                   wrkmemory2, alndatamemory, tfmmemory, alnsmemory, globvarsbuf).Run();
        auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();

        const size_t np = nqystrs * ndbCstrs;
        std::memcpy(alndata, alndatamemory, np * nTDP2OutputAlnData * sizeof(float));
        std::memcpy(tfms, tfmmemory, np * nTTranformMatrix * sizeof(float));
        if (alns) std::memcpy(alns, alnsmemory, std::min(alnsize, szalns));
        return 0;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
