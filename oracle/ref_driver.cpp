// ref_driver.cpp -- runs the reference's own CPU path (libmymp, compiled from
// /root/reference/src by oracle/Makefile) on a query chunk and a reference
// chunk, and returns its result sections.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle and as the timed CPU
// baseline; it is never part of the product.  The sequence below follows
// MpBatch::ProcessBlock (src/libmymp/mpbatch/MpBatch.cpp:283-485) with the
// sequence-similarity pre-filter, the frg3 search (--no-detailed-search) and
// the finalizer thread left out.  After stage 1 the reference's own
// compaction runs (MpReform, PMBatchStrData::FilterStructs): references that
// no query of the block keeps leave it and report nothing
// (gta_ref_last_pass tells which pairs stayed).

#include <omp.h>
#include <chrono>
#include <algorithm>
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
#include "libmymp/mpstage1/MpStageSSRR.h"
#include "libmymp/mpfilter/MpReform.h"
#include "libutil/alpha.h"
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

// per pair of the last gta_ref_process_block call: 1 when the pair's
// reference stayed in its block after the stage-1 filter
static std::vector<unsigned char> g_pass;
// diagnostic: when set, packing a block also moves each kept pair's per-slot
// transforms and per-slot records (slots 1..) to its new index, which the
// reference's packing does not do (MpReform.cpp:57-124 moves slot 0 of the
// records and the output transform only)
static int g_move_slots = 0;

// Moves rows [(q*maxnsteps + si) * n + r] of `width` floats from the
// unpacked layout (n = ndb) to the packed one (n = ndb2, index newndx[r]).
static void move_slot_rows(float* buf, size_t nq, size_t maxnsteps, size_t ndb, size_t ndb2,
                           const std::vector<int>& newndx, size_t width, size_t si0) {
    std::vector<float> old(buf, buf + nq * maxnsteps * ndb * width);
    for (size_t q = 0; q < nq; q++)
        for (size_t si = si0; si < maxnsteps; si++)
            for (size_t r = 0; r < ndb; r++) {
                if (newndx[r] < 0) continue;
                std::memcpy(buf + ((q * maxnsteps + si) * ndb2 + newndx[r]) * width,
                            old.data() + ((q * maxnsteps + si) * ndb + r) * width, width * sizeof(float));
            }
}

// Runs one reference sub-chunk (which the reference's memory layout must hold).
// pass[ql * nstrs + rl] receives 1 for the pairs of references that stay in
// the block after the stage-1 filter; the result sections of the others are
// left untouched (the reference reports nothing for them).
static int run_subchunk(const gta_chunk_view* q, const gta_chunk_view* r, const gta_options* o,
                        MpGlobalMemory* gmem, float* alndata, float* tfms, char* alns, unsigned char* pass) {
    ChunkFields qf, rf;
    fill(q, qf);
    fill(r, rf);

    const size_t nqystrs = q->nstrs, ndbCstrs = r->nstrs;
    const size_t nqyposs = q->nposs, ndbCposs = r->nposs;
    const size_t qystr1len = q->len[0], dbstr1len = r->len[0];
    const size_t qystrnlen = q->len[nqystrs - 1], dbstrnlen = r->len[ndbCstrs - 1];
    const size_t dbxpad = ALIGN_UP(ndbCposs, CUL2CLINESIZE) - ndbCposs;
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

    MpStage1 stg1(2, maxnsteps, minfraglen, o->prescore, stepinit, qb, qe, rb, re,
                  nqystrs, ndbCstrs, nqyposs, ndbCposs, qystr1len, dbstr1len, qystrnlen, dbstrnlen,
                  dbxpad, scores, tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer, maxcoordsbuf,
                  btckdata, wrkmemory, wrkmemoryccd, wrkmemorytm, wrkmemorytmibest, auxwrkmemory,
                  wrkmemory2, alndatamemory, tfmmemory, globvarsbuf);
    stg1.Preinitialize1(false);
    MpSecStr(tmpdpdiagbuffers, qb, qe, rb, re, nqystrs, ndbCstrs, nqyposs, ndbCposs, qystr1len,
             dbstr1len, qystrnlen, dbstrnlen, dbxpad).Run();
    stg1.Run();

    // MpBatch::FilteroutReferences after the stage-1 prescreen
    // (MpBatch.cpp:104-193,371-378): references flagged low-scoring for every
    // query of the block leave it, the others are packed and their pairs'
    // flags reset (MpReform.cpp:28-124)
    size_t ndbCposs2 = ndbCposs, ndbCstrs2 = ndbCstrs, dbstr1len2 = dbstr1len, dbstrnlen2 = dbstrnlen;
    size_t dbxpad2 = dbxpad;
    std::vector<int> newndx(ndbCstrs);
    for (size_t i = 0; i < ndbCstrs; i++) newndx[i] = (int)i;
    if (0.0f < o->prescore) {
        std::vector<uint> filterdata(ndbCstrs * nTFilterData, 0);
        std::vector<const char*> desc(ndbCstrs, "");
        MpReform refm(maxnsteps, qb, qe, rb, re, nullptr, nullptr, nullptr, nullptr, nqystrs, ndbCstrs, nqyposs,
                      ndbCposs, tmpdpdiagbuffers, auxwrkmemory, tfmmemory, globvarsbuf, filterdata.data());
        refm.MakeDbCandidateList();
        for (size_t i = 0; i < ndbCstrs; i++) {
            const uint nd = filterdata[ndbCstrs * fdNewReferenceIndex + i];
            newndx[i] = nd ? (int)nd - 1 : -1;
        }
        PMBatchStrData::FilterStructs(desc.data(), rb, re, filterdata.data());
        ndbCposs2 = PMBatchStrData::GetNoPosits(rb, re);
        ndbCstrs2 = PMBatchStrData::GetNoStructs(rb, re);
        if (ndbCstrs2) {
            dbstr1len2 = PMBatchStrData::GetLengthAt(rb, 0);
            dbstrnlen2 = PMBatchStrData::GetLengthAt(rb, ndbCstrs2 - 1);
        }
        dbxpad2 = ALIGN_UP(ndbCposs2, CUL2CLINESIZE) - ndbCposs2;
        if (ndbCstrs2 && ndbCposs2 && g_move_slots) {
            move_slot_rows(wrkmemorytmibest, nqystrs, maxnsteps, ndbCstrs, ndbCstrs2, newndx, nTTranformMatrix, 0);
            move_slot_rows(wrkmemorytm, nqystrs, maxnsteps, ndbCstrs, ndbCstrs2, newndx, nTTranformMatrix, 0);
            // per-slot records, field-major: [(q*maxnsteps + si) * NF + f] * n + r
            std::vector<float> old(auxwrkmemory, auxwrkmemory + nqystrs * maxnsteps * nTAuxWorkingMemoryVars * ndbCstrs);
            for (size_t q = 0; q < nqystrs; q++)
                for (size_t si = 1; si < maxnsteps; si++)
                    for (int f = 0; f < nTAuxWorkingMemoryVars; f++)
                        for (size_t r = 0; r < ndbCstrs; r++)
                            if (newndx[r] >= 0)
                                auxwrkmemory[((q * maxnsteps + si) * nTAuxWorkingMemoryVars + f) * ndbCstrs2 + newndx[r]] =
                                    old[((q * maxnsteps + si) * nTAuxWorkingMemoryVars + f) * ndbCstrs + r];
        }
        if (ndbCstrs2 && ndbCposs2) refm.SelectAndReformat((int)ndbCstrs2, (int)gmem->GetCurrentMaxDbPos());
    }
    for (size_t ql = 0; ql < nqystrs; ql++)
        for (size_t i = 0; i < ndbCstrs; i++) pass[ql * ndbCstrs + i] = newndx[i] >= 0;
    if (!ndbCstrs2 || !ndbCposs2) return 0;

    if (o->addsearchbyss) {
        // MpBatch.cpp:421-436
        MpStageSSRR stgSSRR(2, maxnsteps, minfraglen, o->prescore, stepinit, qb, qe, rb, re,
                            nqystrs, ndbCstrs2, nqyposs, ndbCposs2, qystr1len, dbstr1len2, qystrnlen, dbstrnlen2,
                            dbxpad2, scores, tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer, maxcoordsbuf,
                            btckdata, wrkmemory, wrkmemoryccd, wrkmemorytm, wrkmemorytmibest, auxwrkmemory,
                            wrkmemory2, alndatamemory, tfmmemory, globvarsbuf);
        stgSSRR.RunSpecialized<true>(2);
    }
    MpStage2 stg2(2, maxnsteps, minfraglen, o->prescore, stepinit, qb, qe, rb, re,
                  nqystrs, ndbCstrs2, nqyposs, ndbCposs2, qystr1len, dbstr1len2, qystrnlen, dbstrnlen2,
                  dbxpad2, scores, tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer, maxcoordsbuf,
                  btckdata, wrkmemory, wrkmemoryccd, wrkmemorytm, wrkmemorytmibest, auxwrkmemory,
                  wrkmemory2, alndatamemory, tfmmemory, globvarsbuf);
    stg2.RunSpecialized<false, true, D02IND_DPSCAN>(2, false, o->scorethld);
    stg2.RunSpecialized<true, false, D02IND_SEARCH>(o->convergence, true, o->scorethld);
    const size_t szalns2 = gmem->GetSizeOfDP2Alns(nqystrs, nqyposs, ndbCposs2, ndbCstrs2);
    std::memset(alnsmemory, 0, szalns2);
    MpStageFin(EQUIVALENCE_DISTANCE * EQUIVALENCE_DISTANCE, maxnsteps, minfraglen, qb, qe, rb, re,
               nqystrs, ndbCstrs2, nqyposs, ndbCposs2, qystr1len, dbstr1len2, qystrnlen, dbstrnlen2,
               dbxpad2, scores, tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer, maxcoordsbuf,
               btckdata, wrkmemory, wrkmemoryccd, wrkmemorytm, wrkmemorytmibest, auxwrkmemory,
               wrkmemory2, alndatamemory, tfmmemory, alnsmemory, globvarsbuf).Run();

    // results of the packed block back at the references' places in the
    // caller's sub-chunk (alignment rows: production_match2aln.cu:130-133)
    std::vector<size_t> dst2(ndbCstrs2);
    for (size_t i = 0; i < ndbCstrs2; i++) dst2[i] = PMBatchStrData::GetAddressAt(rb, i);
    for (size_t ql = 0; ql < nqystrs; ql++) {
        const size_t qlen = q->len[ql];
        for (size_t i = 0; i < ndbCstrs; i++) {
            const int k = newndx[i];
            if (k < 0) continue;
            std::memcpy(alndata + (ql * ndbCstrs + i) * nTDP2OutputAlnData,
                        alndatamemory + (ql * ndbCstrs2 + k) * nTDP2OutputAlnData, nTDP2OutputAlnData * sizeof(float));
            std::memcpy(tfms + (ql * ndbCstrs + i) * nTTranformMatrix,
                        tfmmemory + (ql * ndbCstrs2 + k) * nTTranformMatrix, nTTranformMatrix * sizeof(float));
            const size_t seg = qlen + r->len[i] + 1;
            const size_t sofff = (ql * ndbCposs2 + ndbCstrs2 * (q->dst[ql] + ql)) * 5;
            const size_t slen = ndbCposs2 + ndbCstrs2 * (qlen + 1);
            const size_t sbeg = dst2[k] + (size_t)k * (qlen + 1);
            const size_t dofff = (ql * ndbCposs + ndbCstrs * (q->dst[ql] + ql)) * 5;
            const size_t dlen = ndbCposs + ndbCstrs * (qlen + 1);
            const size_t dbeg = (size_t)r->dst[i] + i * (qlen + 1);
            for (int f = 0; f < 5; f++)
                std::memcpy(alns + dofff + dbeg + dlen * f, alnsmemory + sofff + sbeg + slen * f, seg);
        }
    }
    return 0;
}

// Options are read once per process by the reference (function-local
// statics), so a process must use one option set and one thread count.
// References are processed in sub-chunks that fit the reference's memory
// layout, as its task scheduler does; results are returned for the whole
// reference chunk in the layout of the caller's chunk.
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

        static MpGlobalMemory* gmem = nullptr;
        if (!gmem) gmem = new MpGlobalMemory((size_t)8 * ONEG, 1);
        // the reference's working memory holds at most GetMaxQueriesPerChunk()
        // (--dev-queries-per-chunk, default 2) queries; group them as its
        // scheduler does
        const int qgroup = (int)CuMemoryBase::GetMaxQueriesPerChunk();
        const int nq = q->nstrs, nr = r->nstrs;
        std::vector<float> sa, st;
        std::vector<char> sl;
        if (alns) std::memset(alns, 0, alnsize);
        g_pass.assign((size_t)nq * nr, 0);
        double total = 0.0;
        for (int q0 = 0; q0 < nq; q0 += qgroup) {
            const int q1 = std::min(nq, q0 + qgroup);
            gta_chunk_view qs = *q;
            const unsigned qbase = q->dst[q0];
            std::vector<unsigned> qdst(q1 - q0);
            int qposs = 0;
            for (int i = q0; i < q1; i++) { qdst[i - q0] = q->dst[i] - qbase; qposs += q->len[i]; }
            qs.nstrs = q1 - q0;
            qs.nposs = qposs;
            qs.len = q->len + q0;
            qs.dst = qdst.data();
            qs.type = q->type ? q->type + q0 : nullptr;
            qs.x = q->x + qbase; qs.y = q->y + qbase; qs.z = q->z + qbase;
            qs.ins_ch_ord = q->ins_ch_ord ? q->ins_ch_ord + qbase : nullptr;
            qs.resnum = q->resnum ? q->resnum + qbase : nullptr;
            qs.rsd = q->rsd + qbase;
            qs.ss = q->ss ? q->ss + qbase : nullptr;
            gmem->CalcMaxDbDataChunkSize(qposs);
            const size_t maxpos = gmem->GetCurrentMaxDbPos();
            const size_t maxstrs = gmem->GetCurrentMaxNDbStrs();
            const int nqs = qs.nstrs;
            for (int r0 = 0; r0 < nr;) {
                // greedy sub-chunk [r0, r1) within the memory layout's limits
                int r1 = r0;
                size_t pos = 0;
                while (r1 < nr && (size_t)(r1 - r0) < maxstrs) {
                    size_t np = pos + r->len[r1];
                    if (ALIGN_UP(np, CUL2CLINESIZE) > maxpos) break;
                    pos = np;
                    r1++;
                }
                if (r1 == r0) return -4;
                gta_chunk_view sub = *r;
                const unsigned base = r->dst[r0];
                std::vector<unsigned> sdst(r1 - r0);
                for (int i = r0; i < r1; i++) sdst[i - r0] = r->dst[i] - base;
                sub.nstrs = r1 - r0;
                sub.nposs = (int)pos;
                sub.len = r->len + r0;
                sub.dst = sdst.data();
                sub.type = r->type ? r->type + r0 : nullptr;
                sub.x = r->x + base; sub.y = r->y + base; sub.z = r->z + base;
                sub.ins_ch_ord = r->ins_ch_ord ? r->ins_ch_ord + base : nullptr;
                sub.resnum = r->resnum ? r->resnum + base : nullptr;
                sub.rsd = r->rsd + base;
                sub.ss = r->ss ? r->ss + base : nullptr;
                const size_t nps = (size_t)nqs * sub.nstrs;
                sa.assign(nps * nTDP2OutputAlnData, 0.f);
                st.assign(nps * nTTranformMatrix, 0.f);
                sl.assign(gmem->GetSizeOfDP2Alns((size_t)nqs, (size_t)qs.nposs, (size_t)sub.nposs,
                                                 (size_t)sub.nstrs), 0);
                auto t0 = std::chrono::steady_clock::now();
                std::vector<unsigned char> sp(nps, 0);
                int rc = run_subchunk(&qs, &sub, o, gmem, sa.data(), st.data(), sl.data(), sp.data());
                total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (rc) return rc;
                for (int qi = q0; qi < q1; qi++) {
                    const int ql = qi - q0;
                    for (int ri = r0; ri < r1; ri++) {
                        const int rl = ri - r0;
                        g_pass[(size_t)qi * nr + ri] = sp[(size_t)ql * sub.nstrs + rl];
                        std::memcpy(alndata + ((size_t)qi * nr + ri) * nTDP2OutputAlnData,
                                    sa.data() + ((size_t)ql * sub.nstrs + rl) * nTDP2OutputAlnData,
                                    nTDP2OutputAlnData * sizeof(float));
                        std::memcpy(tfms + ((size_t)qi * nr + ri) * nTTranformMatrix,
                                    st.data() + ((size_t)ql * sub.nstrs + rl) * nTTranformMatrix,
                                    nTTranformMatrix * sizeof(float));
                        if (!alns) continue;
                        // alignment rows: production_match2aln.cu:130-133 layout in both chunks
                        const size_t qlen = q->len[qi];
                        const size_t seg = qlen + r->len[ri] + 1;
                        const size_t sofff = ((size_t)ql * sub.nposs + (size_t)sub.nstrs * (qdst[ql] + ql)) * 5;
                        const size_t slen = (size_t)sub.nposs + (size_t)sub.nstrs * (qlen + 1);
                        const size_t sbeg = (size_t)sdst[rl] + (size_t)rl * (qlen + 1);
                        const size_t dofff = ((size_t)qi * r->nposs + (size_t)nr * (q->dst[qi] + qi)) * 5;
                        const size_t dlen = (size_t)r->nposs + (size_t)nr * (qlen + 1);
                        const size_t dbeg = (size_t)r->dst[ri] + (size_t)ri * (qlen + 1);
                        for (int f = 0; f < 5; f++)
                            std::memcpy(alns + dofff + dbeg + dlen * f, sl.data() + sofff + sbeg + slen * f, seg);
                    }
                }
                r0 = r1;
            }
        }
        if (seconds) *seconds = total;
        return 0;
    } catch (...) {
        return -1;
    }
}

// Diagnostic switch (see g_move_slots); returns the previous setting.
int gta_ref_set_move_slots(int on) {
    const int prev = g_move_slots;
    g_move_slots = on ? 1 : 0;
    return prev;
}

// Per pair of the last gta_ref_process_block call (nq*nr bytes): 1 when the
// pair's reference stayed in its block after the stage-1 filter, 0 when the
// reference left it (no results are reported for that pair).
int gta_ref_last_pass(unsigned char* out, size_t n) {
    if (n != g_pass.size()) return -1;
    std::memcpy(out, g_pass.data(), n);
    return 0;
}

// Secondary structure of every position of one chunk on the reference's
// MpSecStr (protein rule and nucleic-acid base pairing, MpSecStr.h:50-56),
// the chunk serving as both query and reference chunk; out receives nposs
// letters; pos (may be NULL) receives the pairing partner the reference
// recorded for each position (-1: none).
int gta_ref_secstr(const gta_chunk_view* c, char* out, int* pos) {
    try {
        ChunkFields qf, rf;
        fill(c, qf);
        fill(c, rf);
        const size_t ns = c->nstrs, np = c->nposs;
        const size_t L1 = c->len[0], Ln = c->len[ns - 1];
        std::vector<float> tmp(3 * np + 64, 0.f);
        MpSecStr(tmp.data(), qf.beg, qf.end, rf.beg, rf.end, ns, ns, np, np, L1, L1, Ln, Ln,
                 ALIGN_UP(np, CUL2CLINESIZE) - np).Run();
        std::memcpy(out, rf.beg[pmv2Dss], np);
        if (pos)
            for (size_t i = 0; i < np; i++) pos[i] = (int)tmp[i];
        return 0;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"

// One DP fill + BtckToMatched of one pair on the reference's code: the
// superposition fill ExecDPwBtck128xKernel (ssrr 0) or the fill from SS and
// residue similarity ExecDPSSwBtck128xKernel<true> (ssrr 1, as
// MpStageSSRR::Align calls it, MpStageSSRR.h:94-120).
static int ref_dp_pair(const gta_chunk_view& q, const gta_chunk_view& r, int ssrr, const float* tfm, int gap0,
                       float gapopen, float wss, float wrr, unsigned char* bt, int* qpos, int* rpos) {
    const int Lq = q.nposs, Lr = r.nposs;
    const float *qx = q.x, *rx = r.x;
    {
        ChunkFields qf, rf;
        fill(&q, qf);
        fill(&r, rf);
        static MpGlobalMemory* gmem = nullptr;
        if (!gmem) gmem = new MpGlobalMemory((size_t)2 * ONEG, 1);
        gmem->CalcMaxDbDataChunkSize(Lq);
        const size_t dbxpad = ALIGN_UP((size_t)Lr, CUL2CLINESIZE) - Lr;
        if (gmem->GetCurrentMaxDbPos() < Lr + dbxpad) return -4;
        auto sec = [&](int s) { return gmem->GetHeap() + gmem->GetHeapSectionOffset(0, s); };
        const uint maxnsteps = MpGlobalMemory::GetMaxNFragSteps();
        const uint minfraglen = GetMinFragLengthForAln(MpGlobalMemory::GetMaxAlnLength());
        float* tmpdpdiagbuffers = (float*)sec(MpGlobalMemory::ddsBegOfDPDiagScores);
        float* tmpdpbotbuffer = (float*)sec(MpGlobalMemory::ddsBegOfDPBottomScores);
        float* tmpdpalnpossbuffer = (float*)sec(MpGlobalMemory::ddsBegOfDPAlignedPoss);
        char* btckdata = sec(MpGlobalMemory::ddsBegOfDPBackTckData);
        float* wrkmemorytmibest = (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryTMibest);
        float* auxwrkmemory = (float*)sec(MpGlobalMemory::ddsBegOfAuxWrkMemory);
        float* tfmmemory = (float*)sec(MpGlobalMemory::ddsBegOfTfmMatrices);
        float* alndatamemory = (float*)sec(MpGlobalMemory::ddsBegOfDP2AlnData);
        MpStage1 stg1(2, maxnsteps, minfraglen, 0.f, 1, qf.beg, qf.end, rf.beg, rf.end, 1, 1, Lq, Lr,
                      Lq, Lr, Lq, Lr, dbxpad, (float*)sec(MpGlobalMemory::ddsBegOfMtxScores),
                      tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer,
                      (uint*)sec(MpGlobalMemory::ddsBegOfDPMaxCoords), btckdata,
                      (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory),
                      (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryCCD),
                      (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryTM), wrkmemorytmibest, auxwrkmemory,
                      (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory2), alndatamemory, tfmmemory,
                      (uint*)sec(MpGlobalMemory::ddsBegOfGlobVars));
        stg1.Preinitialize1(false);
        for (int f = 0; f < nTTranformMatrix; f++) wrkmemorytmibest[f] = tfm[f];
        MpDPHub dph(maxnsteps, qf.beg, qf.end, rf.beg, rf.end, 1, 1, Lq, Lr, Lq, Lr, dbxpad,
                    tmpdpdiagbuffers, tmpdpbotbuffer, tmpdpalnpossbuffer,
                    (uint*)sec(MpGlobalMemory::ddsBegOfDPMaxCoords), btckdata,
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory),
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryCCD),
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryTM), wrkmemorytmibest, auxwrkmemory,
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory2), alndatamemory, tfmmemory,
                    (uint*)sec(MpGlobalMemory::ddsBegOfGlobVars));
        if (ssrr)
            dph.ExecDPSSwBtck128xKernel<true>(gapopen, wss, wrr, qf.beg, rf.beg, auxwrkmemory, tmpdpdiagbuffers,
                                              tmpdpbotbuffer, btckdata);
        else if (gap0)
            dph.ExecDPwBtck128xKernel<false, false, true, D02IND_SEARCH, false>(
                gapopen, 0, qf.beg, rf.beg, wrkmemorytmibest, auxwrkmemory, tmpdpdiagbuffers,
                tmpdpbotbuffer, btckdata);
        else
            dph.ExecDPwBtck128xKernel<false, false, false, D02IND_SEARCH, false>(
                gapopen, 0, qf.beg, rf.beg, wrkmemorytmibest, auxwrkmemory, tmpdpdiagbuffers,
                tmpdpbotbuffer, btckdata);
        dph.BtckToMatched128xKernel<false, false>(0, qf.beg, rf.beg, btckdata, auxwrkmemory,
                                                  tmpdpalnpossbuffer);
        const size_t dblen = Lr + dbxpad;
        if (bt)
            for (int i = 0; i < Lq; i++)
                for (int j = 0; j < Lr; j++) bt[(size_t)i * Lr + j] = (unsigned char)btckdata[(size_t)i * dblen + j];
        const int n = (int)auxwrkmemory[tawmvNAlnPoss];
        // matched coordinates identify the positions (the walk order is kept)
        for (int k = 0; k < n; k++) {
            const float x = tmpdpalnpossbuffer[k + dpapsQRYx * dblen];
            const float y = tmpdpalnpossbuffer[k + dpapsRFNx * dblen];
            int a = -1, b = -1;
            for (int i = 0; i < Lq && a < 0; i++) if (qx[i] == x) a = i;
            for (int j = 0; j < Lr && b < 0; j++) if (rx[j] == y) b = j;
            qpos[k] = a;
            rpos[k] = b;
        }
        return n;
    }
}

extern "C" {

// The reference's DP fill + backtrace (MpDPHub::ExecDPwBtck128xKernel and
// BtckToMatched128xKernel, MpDPHub.cpp:73-388,1500-1610) for one pair under a
// given transform; bt receives the Lq x Lr backtrace codes and qpos/rpos the
// matched positions in walk order.  Returns the number of matched pairs.
int gta_ref_dp(const float* qx, const float* qy, const float* qz, int Lq, const float* rx,
               const float* ry, const float* rz, int Lr, const float* tfm, int gap0, float gapopen,
               unsigned char* bt, int* qpos, int* rpos) {
    try {
        std::vector<int> ql{Lq}, rl{Lr}, zq(Lq, 0), zr(Lr, 0);
        std::vector<unsigned> d0{0};
        std::vector<char> aq(Lq, 'A'), ar(Lr, 'A');
        gta_chunk_view q{1, Lq, ql.data(), d0.data(), nullptr, qx, qy, qz, zq.data(), nullptr, aq.data(), nullptr};
        gta_chunk_view r{1, Lr, rl.data(), d0.data(), nullptr, rx, ry, rz, zr.data(), nullptr, ar.data(), nullptr};
        return ref_dp_pair(q, r, 0, tfm, gap0, gapopen, 0.f, 0.f, bt, qpos, rpos);
    } catch (...) {
        return -1;
    }
}

// The SS/residue fill for one pair given its SS letters and residue codes
// (gap open gapopen, weights wss and wrr); matched positions in walk order.
int gta_ref_dp_ssrr(const char* qss, const char* qrsd, int Lq, const char* rss, const char* rrsd, int Lr,
                    float gapopen, float wss, float wrr, unsigned char* bt, int* qpos, int* rpos) {
    try {
        std::vector<int> ql{Lq}, rl{Lr}, zq(Lq, 0), zr(Lr, 0);
        std::vector<unsigned> d0{0};
        // distinct x coordinates identify the matched positions
        std::vector<float> qx(Lq), rx(Lr), qz(Lq, 0.f), rz(Lr, 0.f);
        for (int i = 0; i < Lq; i++) qx[i] = (float)i;
        for (int j = 0; j < Lr; j++) rx[j] = (float)j;
        const float tfm[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
        gta_chunk_view q{1, Lq, ql.data(), d0.data(), nullptr, qx.data(), qz.data(), qz.data(), zq.data(),
                         nullptr, qrsd, qss};
        gta_chunk_view r{1, Lr, rl.data(), d0.data(), nullptr, rx.data(), rz.data(), rz.data(), zr.data(),
                         nullptr, rrsd, rss};
        return ref_dp_pair(q, r, 1, tfm, 0, gapopen, wss, wrr, bt, qpos, rpos);
    } catch (...) {
        return -1;
    }
}

// The reference's score-only fill, MpDPHub::ExecDPScore128xKernel<GAP0=true>
// (MpDPHub.cpp:389-647), for K superpositions of one pair (tfms[k*12]); each
// score is read where MpStageFrg3::SortBestDPscoresAndTMsAmongDPswifts reads it
// (MpStageFrg3.h:1180-1191).  Returns 0, or < 0 on error.
int gta_ref_dpscore(const float* qx, const float* qy, const float* qz, int Lq, const float* rx,
                    const float* ry, const float* rz, int Lr, const float* tfms, int K, float* scores) {
    try {
        std::vector<int> ql{Lq}, rl{Lr}, zq(Lq, 0), zr(Lr, 0);
        std::vector<unsigned> d0{0};
        std::vector<char> aq(Lq, 'A'), ar(Lr, 'A');
        gta_chunk_view q{1, Lq, ql.data(), d0.data(), nullptr, qx, qy, qz, zq.data(), nullptr, aq.data(), nullptr};
        gta_chunk_view r{1, Lr, rl.data(), d0.data(), nullptr, rx, ry, rz, zr.data(), nullptr, ar.data(), nullptr};
        ChunkFields qf, rf;
        fill(&q, qf);
        fill(&r, rf);
        static MpGlobalMemory* gmem = nullptr;
        if (!gmem) gmem = new MpGlobalMemory((size_t)2 * ONEG, 1);
        gmem->CalcMaxDbDataChunkSize(Lq);
        const size_t dbxpad = ALIGN_UP((size_t)Lr, CUL2CLINESIZE) - Lr;
        if (gmem->GetCurrentMaxDbPos() < Lr + dbxpad) return -4;
        auto sec = [&](int s) { return gmem->GetHeap() + gmem->GetHeapSectionOffset(0, s); };
        const uint maxnsteps = MpGlobalMemory::GetMaxNFragSteps();
        const int depth = CLOptions::GetC_DEPTH();
        const int nconfigs = CUS1_TBSP_DPSCORE_TOP_N *
                             ((depth <= CLOptions::csdDeep) ? 3 : ((depth <= CLOptions::csdHigh) ? 2 : 1));
        if (K < 1 || K > nconfigs || (uint)nconfigs > maxnsteps) return -2;
        float* tmpdpdiagbuffers = (float*)sec(MpGlobalMemory::ddsBegOfDPDiagScores);
        float* tmpdpbotbuffer = (float*)sec(MpGlobalMemory::ddsBegOfDPBottomScores);
        float* wrkmemtm = (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryTM);
        float* auxwrkmemory = (float*)sec(MpGlobalMemory::ddsBegOfAuxWrkMemory);
        // one query and one reference: slot si's transform at si*12, its
        // convergence flag at (si*nTAuxWorkingMemoryVars + tawmvConverged)
        for (int si = 0; si < nconfigs; si++) {
            for (int f = 0; f < nTTranformMatrix; f++)
                wrkmemtm[si * nTTranformMatrix + f] = si < K ? tfms[si * 12 + f] : ((f == 0 || f == 4 || f == 8) ? 1.f : 0.f);
            auxwrkmemory[si * nTAuxWorkingMemoryVars + tawmvConverged] = 0.f;
        }
        MpDPHub dph(maxnsteps, qf.beg, qf.end, rf.beg, rf.end, 1, 1, Lq, Lr, Lq, Lr, dbxpad,
                    tmpdpdiagbuffers, tmpdpbotbuffer, (float*)sec(MpGlobalMemory::ddsBegOfDPAlignedPoss),
                    (uint*)sec(MpGlobalMemory::ddsBegOfDPMaxCoords), sec(MpGlobalMemory::ddsBegOfDPBackTckData),
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory),
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryCCD), wrkmemtm,
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemoryTMibest), auxwrkmemory,
                    (float*)sec(MpGlobalMemory::ddsBegOfWrkMemory2),
                    (float*)sec(MpGlobalMemory::ddsBegOfDP2AlnData), (float*)sec(MpGlobalMemory::ddsBegOfTfmMatrices),
                    (uint*)sec(MpGlobalMemory::ddsBegOfGlobVars));
        {
            // no value of an earlier call may survive in the buffers read below
            const size_t dblen0 = (size_t)Lr + dbxpad;
            std::fill(tmpdpdiagbuffers, tmpdpdiagbuffers + (size_t)nTDPDiagScoreSections * nTDPDiagScoreSubsections *
                                                               (size_t)nconfigs * dblen0, 0.f);
            std::fill(tmpdpbotbuffer, tmpdpbotbuffer + (size_t)nTDPDiagScoreSubsections * (size_t)nconfigs * dblen0,
                      0.f);
        }
        dph.ExecDPScore128xKernel<true>(0.0f, qf.beg, rf.beg, wrkmemtm, auxwrkmemory, tmpdpdiagbuffers,
                                        tmpdpbotbuffer);
        const int dblen = Lr + (int)dbxpad;
        for (int si = 0; si < K; si++) {
            const int yofff = si * dblen;
            const int doffs = nTDPDiagScoreSections * nTDPDiagScoreSubsections * yofff;
            scores[si] = tmpdpdiagbuffers[doffs + 0 + Lr - 1];
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

// The reference's pairwise residue scores (GONNET_SCORES, alpha.cpp:198-206):
// 26 x 26 floats.
void gta_ref_gonnet(float* out) { std::memcpy(out, GONNET_SCORES.data_, sizeof(GONNET_SCORES.data_)); }

}  // extern "C"

