// gta_capi.hip -- implementation of the C ABI (include/gta/gtalign_amd.h).
//
// Host orchestration of the stage on one HIP stream: the sequence of
// MpBatch::ProcessBlock / CuBatch::ProcessBlock (MpBatch.cpp:283-485,
// CuBatch.cu:464-681) -- stage 1, stage 2 in two passes, final stage -- with
// every per-pair quantity resident in HBM between kernels.  The sequence-
// similarity pre-filter and the frg3 search are off, their defaults for this
// stage; pairs failing the pre-score test are flagged and skipped by every
// later kernel instead of being compacted out of the chunk.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gta/gtalign_amd.h"
#include "gta/gtalign_amd_testing.h"
#include "gta_launch.h"

using namespace gta;

namespace {

thread_local std::string g_tls_error;

// Bump allocator over the device scratch of gta_process_noalloc: while one is
// set for the calling thread, device buffers are carved from it (256-byte
// aligned) instead of being allocated, and never freed.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, off = 0;
    bool over = false;
};
thread_local Arena* g_arena = nullptr;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        const size_t need = std::max<size_t>(n, 256);
        if (g_arena) {
            const size_t at = (g_arena->off + 255) & ~(size_t)255;
            if (at > g_arena->cap || need > g_arena->cap - at) {
                g_arena->over = true;
                return hipErrorOutOfMemory;
            }
            p = g_arena->base + at;
            bytes = need;
            g_arena->off = at + need;
            return hipSuccess;
        }
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    void release() {
        if (p && !g_arena) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return (T*)p; }
};

struct ChunkDev {
    int nstrs = 0, nposs = 0;
    std::vector<int> len;
    std::vector<unsigned> dst;
    std::vector<int> na;
    DevBuf b_len, b_dst, b_na, b_x, b_y, b_z, b_rsd, b_ss, b_napos;
    DevChunk view() const {
        DevChunk c;
        c.nstrs = nstrs; c.nposs = nposs;
        c.len = b_len.as<int>(); c.dst = b_dst.as<unsigned>(); c.na = b_na.as<int>();
        c.x = b_x.as<float>(); c.y = b_y.as<float>(); c.z = b_z.as<float>();
        c.rsd = b_rsd.as<char>(); c.ss = b_ss.as<char>(); c.napos = b_napos.as<int>();
        return c;
    }
    void release() {
        b_len.release(); b_dst.release(); b_na.release(); b_x.release(); b_y.release();
        b_z.release(); b_rsd.release(); b_ss.release(); b_napos.release();
    }
};

int GetMaxNFragSteps(int maxalnlen, int step, int fraglen) {
    return (int)ceilf((float)(maxalnlen + step - fraglen) / step);
}

}  // namespace

// work-queue counters of the batched form (one per persistent launch)
#define GTA_MAX_QUEUES 128

struct gta_ctx {
    int device = 0;
    void* scratch = nullptr;    // this device's address of `scratch` (gta_process_noalloc)
    gta_options opts;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int max_qrys = 0, max_qposs = 0, max_refs = 0, max_rposs = 0;
    size_t lds_max = 65536;
    ChunkDev q, r;
    bool have_q = false, have_r = false;
    // pair state
    DevBuf s_flags, s_grand, s_best0, s_prev, s_naln, s_qpos, s_rpos, s_fndx, s_fpos;
    DevBuf s_tibest, s_tgrand, s_out, s_aln, s_btck, s_btckoff, s_rowbuf, s_cells, s_alns, s_algb;
    // DP size buckets: pairs filled by W waves (and their ring rows)
    DevBuf s_dpsmall, s_dpbig, s_ring, s_ringoff, s_fail, s_gstrip0, s_gsync;
    DpBuckets dpb{};
    // a pair is filled by several waves when Lq*Lr >= mw_min_cells and it
    // has at least mw_min_strips 64-row strips (and Lr >= 256)
    long long mw_min_cells = 1ll << 22;
    int mw_min_strips = 16;
    // those pairs are pipelined over the whole device instead when they number
    // at most half the CUs (dp_grid 0), always (1) or never (2)
    int dp_grid = 0;
    // per pair: DP launch form (DPF_*) of the resident chunks
    std::vector<unsigned char> dpform;
    // DP launch-form autotuning: on/off, cache file, device tag and the
    // chosen form per length bucket ("strip class, Lr class, count class")
    int autotune = 0;
    std::string tune_path, tune_tag;
    std::map<std::string, int> tune;
    double tune_seconds = 0.0;
    // refinement split: pairs whose shorter structure has at least
    // split_min_len residues spread their slots over several workgroups
    DevBuf s_rsmall, s_rbig, s_parts;
    RefSplit rsp{};
    int split_min_len = 2048;
    // per part: 4 seed diagonals per 4-lane group of the seed search and one
    // fragment slot per 8-lane group of the refinements
    int split_seeds = 64 * SEARCH_WAVES, split_slots = 8 * REFINE_WAVES;
    // batched form: live-pair work lists, queue counters, the captured graph
    DevBuf s_bigmask, s_lista, s_listb, s_listc, s_lcnt, s_ltot, s_heads;
    DevBuf s_sctfm, s_scflag, s_scout;      // score-only DP test hook
    const unsigned char* bigmask = nullptr;
    int ncu = 256;
    hipGraphExec_t gexec = nullptr;
    bool gvalid = false;
    // measurement
    int profiling = 0;
    std::vector<hipEvent_t> evpool;
    std::vector<std::pair<int, int>> evrec;   // (kernel class, first event index)
    int kc_launches[KC_COUNT] = {0};
    double kc_seconds[KC_COUNT] = {0};
    double kc_bytes[KC_COUNT] = {0};
    double kc_flops[KC_COUNT] = {0};
    double cells_useful = 0, cells_launched = 0;  // DP cells of the last run
    // host results
    std::vector<float> h_out, h_tfm;
    std::vector<char> h_alns;
    std::string err;
};

static int fail(gta_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg; else g_tls_error = msg;
    return code;
}
#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, GTA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" {

void gta_default_options(gta_options* o) {
    if (!o) return;
    o->prescore = 0.4f;
    o->scorethld = 0.5f;
    o->convergence = 18;
    o->depth = GTA_DEPTH_MEDIUM;
    o->refinement = GTA_REF_ONE;
    o->symmetric = 0;
    o->nodeletions = 0;
    o->referenced = 0;
    o->precision = GTA_PREC_FP32;
    o->batched = 0;
    o->addsearchbyss = 0;
    o->qgroup = 0;
}

size_t gta_alignment_section_size(int nq, int nqposs, int nr, int nrposs) {
    return ((size_t)nrposs * nq + (size_t)nr * ((size_t)nqposs + nq)) * GTA_ALN_NROWS;
}

const char* gta_last_error(const gta_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_tls_error.c_str();
}

// Device scratch of gta_process_noalloc: one array per device, held until the
// process ends.  Host code never names it except in hipGetSymbolAddress; it
// uses each device's address (scratch_address) everywhere else.
static __device__ __align__(256) unsigned char scratch[GTA_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::map<int, void*> g_scratch_ptr;

// this device's address of `scratch`; the device must be current
static hipError_t scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    *out = p;
    return hipSuccess;
}

static int check_options(const gta_options* o) {
    if (!(o->prescore >= 0.f && o->prescore < 1.f)) return 0;
    if (!(o->scorethld >= 0.f && o->scorethld < 1.f)) return 0;
    if (o->convergence < 1 || o->convergence > 30) return 0;
    if (o->depth < 0 || o->depth > GTA_DEPTH_SHALLOW) return 0;
    if (o->refinement < 0 || o->refinement > GTA_REF_FULLA) return 0;
    if (o->precision != GTA_PREC_FP32 && o->precision != GTA_PREC_MIXED) return 0;
    if (o->batched != 0 && o->batched != 1) return 0;
    if (o->addsearchbyss != 0 && o->addsearchbyss != 1) return 0;
    if (o->qgroup < 0) return 0;
    return 1;
}

int gta_create(int device, const gta_options* opts, int max_qrys, int max_qry_poss, int max_refs,
               int max_ref_poss, gta_ctx** out) {
    if (!out) return fail(nullptr, GTA_ERR_ARG, "null output pointer");
    *out = nullptr;
    gta_options o;
    gta_default_options(&o);
    if (opts) o = *opts;
    if (!check_options(&o)) return fail(nullptr, GTA_ERR_ARG, "option out of range");
    if (max_qrys < 1 || max_qry_poss < 1 || max_refs < 1 || max_ref_poss < 1)
        return fail(nullptr, GTA_ERR_ARG, "invalid capacity");
    if (max_qrys > 100) return fail(nullptr, GTA_ERR_TOOMANY, "more than 100 queries per chunk");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, GTA_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    e = upload_length_tables();
    if (e != hipSuccess) return fail(nullptr, GTA_ERR_HIP, std::string("length tables: ") + hipGetErrorString(e));
    gta_ctx* c = new gta_ctx();
    c->device = device;
    c->opts = o;
    c->max_qrys = max_qrys; c->max_qposs = max_qry_poss;
    c->max_refs = max_refs; c->max_rposs = max_ref_poss;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        c->lds_max = (size_t)lds;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        c->ncu = ncu;
    e = scratch_address(device, &c->scratch);
    if (e != hipSuccess) {
        delete c;
        return fail(nullptr, GTA_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return fail(nullptr, GTA_ERR_HIP, "stream/event creation failed");
    }
    *out = c;
    return GTA_OK;
}

// every device buffer of a context
static std::vector<DevBuf*> all_bufs(gta_ctx* c) {
    return {&c->q.b_len, &c->q.b_dst, &c->q.b_na, &c->q.b_x, &c->q.b_y, &c->q.b_z, &c->q.b_rsd, &c->q.b_ss,
            &c->q.b_napos, &c->r.b_len, &c->r.b_dst, &c->r.b_na, &c->r.b_x, &c->r.b_y, &c->r.b_z,
            &c->r.b_rsd, &c->r.b_ss, &c->r.b_napos,
            &c->s_flags, &c->s_grand, &c->s_best0, &c->s_prev, &c->s_naln, &c->s_qpos,
            &c->s_rpos, &c->s_fndx, &c->s_fpos, &c->s_tibest, &c->s_tgrand, &c->s_out,
            &c->s_aln, &c->s_btck, &c->s_btckoff, &c->s_rowbuf, &c->s_cells, &c->s_alns,
            &c->s_algb, &c->s_dpsmall, &c->s_dpbig, &c->s_ring, &c->s_ringoff, &c->s_fail,
            &c->s_gstrip0, &c->s_gsync,
            &c->s_rsmall, &c->s_rbig, &c->s_parts, &c->s_bigmask, &c->s_lista, &c->s_listb,
            &c->s_listc, &c->s_lcnt, &c->s_ltot, &c->s_heads, &c->s_sctfm, &c->s_scflag,
            &c->s_scout};
}

void gta_destroy(gta_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (DevBuf* b : all_bufs(c)) b->release();
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int upload_chunk(gta_ctx* c, ChunkDev& d, const gta_chunk_view* v, int maxstrs, int maxposs) {
    if (!v || !v->len || !v->dst || !v->x || !v->y || !v->z || !v->rsd)
        return fail(c, GTA_ERR_ARG, "null chunk field");
    if (v->nstrs < 1 || v->nposs < 1) return fail(c, GTA_ERR_EMPTY, "no structures in the chunk");
    if (v->nstrs > maxstrs || v->nposs > maxposs)
        return fail(c, GTA_ERR_SIZE, "chunk larger than the context capacity");
    long long tot = 0;
    for (int i = 0; i < v->nstrs; i++) {
        if (v->len[i] < 1) return fail(c, GTA_ERR_ARG, "empty structure");
        if (v->len[i] > 65535) return fail(c, GTA_ERR_LENGTH, "structure longer than 65535 residues");
        if (v->dst[i] != (unsigned)tot) return fail(c, GTA_ERR_ARG, "structure offsets are not prefix sums");
        if (i && v->len[i] > v->len[i - 1]) return fail(c, GTA_ERR_ARG, "structures not sorted by descending length");
        tot += v->len[i];
    }
    if (tot != v->nposs) return fail(c, GTA_ERR_ARG, "nposs does not match the lengths");
    // the pair score's division is exact only for operands far below the
    // float exponent limit: coordinates must be finite and |x| <= GTA_MAX_COORD
    for (int i = 0; i < v->nposs; i++) {
        const float a[3] = {v->x[i], v->y[i], v->z[i]};
        for (float f : a)
            if (!(fabsf(f) <= GTA_MAX_COORD))
                return fail(c, GTA_ERR_ARG, "coordinate not finite or above GTA_MAX_COORD in magnitude");
    }
    (void)hipSetDevice(c->device);
    d.nstrs = v->nstrs; d.nposs = v->nposs;
    d.len.assign(v->len, v->len + v->nstrs);
    d.dst.assign(v->dst, v->dst + v->nstrs);
    d.na.assign(v->nstrs, 0);
    if (v->ins_ch_ord)
        for (int i = 0; i < v->nstrs; i++) d.na[i] = v->ins_ch_ord[v->dst[i]] > 0;
    const size_t ns = v->nstrs, np = v->nposs;
    HIPCHK(c, d.b_len.ensure(ns * 4));
    HIPCHK(c, d.b_dst.ensure(ns * 4));
    HIPCHK(c, d.b_na.ensure(ns * 4));
    HIPCHK(c, d.b_x.ensure(np * 4));
    HIPCHK(c, d.b_y.ensure(np * 4));
    HIPCHK(c, d.b_z.ensure(np * 4));
    HIPCHK(c, d.b_rsd.ensure(np));
    HIPCHK(c, d.b_ss.ensure(np));
    HIPCHK(c, d.b_napos.ensure(np * 4));
    HIPCHK(c, hipMemcpy(d.b_len.p, d.len.data(), ns * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.b_dst.p, d.dst.data(), ns * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.b_na.p, d.na.data(), ns * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.b_x.p, v->x, np * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.b_y.p, v->y, np * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.b_z.p, v->z, np * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d.b_rsd.p, v->rsd, np, hipMemcpyHostToDevice));
    // the stage assigns every position's SS itself (k_secstr), as the
    // reference does (MpSecStr::Run, MpSecStr.h:50-56); blank until then
    HIPCHK(c, hipMemset(d.b_ss.p, ' ', np));
    return GTA_OK;
}

int gta_set_queries(gta_ctx* c, const gta_chunk_view* q) {
    if (!c) return fail(nullptr, GTA_ERR_ARG, "null context");
    if (q && q->nstrs > 100) return fail(c, GTA_ERR_TOOMANY, "more than 100 queries per chunk");
    int rc = upload_chunk(c, c->q, q, c->max_qrys, c->max_qposs);
    c->have_q = (rc == GTA_OK);
    c->gvalid = false;
    return rc;
}

int gta_set_references(gta_ctx* c, const gta_chunk_view* r) {
    if (!c) return fail(nullptr, GTA_ERR_ARG, "null context");
    int rc = upload_chunk(c, c->r, r, c->max_refs, c->max_rposs);
    c->have_r = (rc == GTA_OK);
    c->gvalid = false;
    return rc;
}

// grid sizes of the chunk (MpStage1.cpp:284-296,506-516; MpStageFin.cpp:304-309)
static Geom make_geom(const gta_ctx* c) {
    Geom g;
    const int qystr1len = c->q.len.front(), qystrnlen = c->q.len.back();
    const int dbstr1len = c->r.len.front(), dbstrnlen = c->r.len.back();
    const int minlenmin = std::min(qystrnlen, dbstrnlen);
    const int minalnmin = std::max(minlenmin >> 1, 5);
    g.stepinit = (c->opts.depth == GTA_DEPTH_SHALLOW) ? 5 : 1;
    g.n1 = minalnmin - dbstr1len;
    const int n2 = qystr1len - minalnmin;
    g.nseeds = (n2 - g.n1) / g.stepinit + 1;
    const int maxalnmax = std::min(qystr1len, dbstr1len);
    const int minfraglen = 4;  // GetMinFragLengthForAln(GetMaxAlnLength()) for any length >= 64
    g.nlocsteps = GetMaxNFragSteps(maxalnmax, SFRAGSTEP, minfraglen) * NMAXSUBFRAGS;
    g.maxnfragfcts = std::min(2 * maxalnmax, 128);
    return g;
}

// ---- DP launch forms ----------------------------------------------------------
// Every pair's fill runs in one of five forms: one wave (k_dp), one workgroup
// of 4, 8 or 16 waves (k_dp_mw) or the device-wide strip pipeline
// (k_dp_grid).  By default large pairs (Lq*Lr >= mw_min_cells, >= mw_min_strips
// strips, Lr >= 256) take the grid form when they number at most half the CUs
// and the W-wave form otherwise.  With autotuning on, pairs of >= 4 strips and
// >= 2^18 cells are grouped by length bucket (log2 of strips, log2 of Lr,
// log2 of the bucket's pair count) and each bucket's form is the fastest of
// the candidates in a timed fill of the whole chunk, kept per context and in
// the cache file.
enum { DPF_ONE = 0, DPF_W4 = 1, DPF_W8 = 2, DPF_W16 = 3, DPF_GRID = 4, DPF_N = 5 };

static int ilog2ll(long long v) {
    int r = 0;
    while (v > 1) { v >>= 1; r++; }
    return r;
}

// per-pair forms -> device lists [one | W4 | W8 | W16 | grid], ring offsets and
// strip counters
static int upload_dp_buckets(gta_ctx* c, const std::vector<unsigned char>& form, DpBuckets& B) {
    const int nr = c->r.nstrs;
    const size_t np = form.size();
    std::vector<int> lists[DPF_N];
    for (size_t p = 0; p < np; p++) lists[form[p]].push_back((int)p);
    B = DpBuckets{};
    HIPCHK(c, c->s_fail.ensure(4));
    B.failflag = c->s_fail.as<int>();
    if (lists[DPF_ONE].size() == np) return GTA_OK;     // small == nullptr: every pair one wave
    std::vector<int> all;
    std::vector<unsigned long long> offs;
    size_t start[DPF_N];
    unsigned long long ringtot = 0;
    for (int f = 0; f < DPF_N; f++) {
        start[f] = all.size();
        for (int pidx : lists[f]) {
            all.push_back(pidx);
            const int Lq = c->q.len[pidx / nr], Lr = c->r.len[pidx % nr];
            offs.push_back(ringtot);
            if (f == DPF_ONE) continue;
            const unsigned long long rows =
                f == DPF_GRID ? (unsigned long long)((Lq + 63) / 64) : (4ull << (f - DPF_W4));
            ringtot += rows * (unsigned long long)Lr;
        }
    }
    std::vector<int> s0;
    int tot = 0;
    for (int pidx : lists[DPF_GRID]) {
        s0.push_back(tot);
        tot += (c->q.len[pidx / nr] + 63) / 64;
    }
    s0.push_back(tot);
    HIPCHK(c, c->s_dpbig.ensure(all.size() * 4));
    HIPCHK(c, c->s_ringoff.ensure(offs.size() * 8));
    HIPCHK(c, c->s_ring.ensure(std::max<unsigned long long>(ringtot, 1) * 4));
    HIPCHK(c, c->s_gstrip0.ensure(s0.size() * 4));
    HIPCHK(c, c->s_gsync.ensure(((size_t)tot + 1) * 4));
    HIPCHK(c, hipMemcpy(c->s_dpbig.p, all.data(), all.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->s_ringoff.p, offs.data(), offs.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->s_gstrip0.p, s0.data(), s0.size() * 4, hipMemcpyHostToDevice));
    const int* L = c->s_dpbig.as<int>();
    const unsigned long long* O = c->s_ringoff.as<unsigned long long>();
    B.small = L + start[DPF_ONE];
    B.nsmall = (int)lists[DPF_ONE].size();
    for (int g = 0; g < 3; g++) {
        B.mw[g] = L + start[DPF_W4 + g];
        B.nmw[g] = (int)lists[DPF_W4 + g].size();
        B.mwoff[g] = O + start[DPF_W4 + g];
    }
    B.ring = c->s_ring.as<float>();
    B.grid = L + start[DPF_GRID];
    B.ngrid = (int)lists[DPF_GRID].size();
    B.gstrip0 = c->s_gstrip0.as<int>();
    B.gtotal = tot;
    B.gsync = c->s_gsync.as<int>();
    B.growoff = O + start[DPF_GRID];
    B.gworkers = std::max(1, std::min(tot, 2 * c->ncu));
    return GTA_OK;
}

static void default_dp_forms(const gta_ctx* c, std::vector<unsigned char>& form) {
    const int nq = c->q.nstrs, nr = c->r.nstrs;
    std::vector<int> big;
    int minstrips = 1 << 30;
    for (int qi = 0; qi < nq; qi++) {
        const int Lq = c->q.len[qi], ns = (Lq + 63) / 64;
        for (int ri = 0; ri < nr; ri++) {
            const int Lr = c->r.len[ri];
            if ((long long)Lq * Lr >= c->mw_min_cells && ns >= c->mw_min_strips && Lr >= 256) {
                big.push_back(qi * nr + ri);
                minstrips = std::min(minstrips, ns);
            }
        }
    }
    if (big.empty()) return;
    const bool grid = c->dp_grid == 1 || (c->dp_grid == 0 && (int)big.size() * 2 <= c->ncu);
    const int f = grid ? DPF_GRID : (minstrips >= 32 ? DPF_W16 : (minstrips >= 16 ? DPF_W8 : DPF_W4));
    for (int pidx : big) form[pidx] = (unsigned char)f;
}

// time of one DP fill of the whole chunk under the given forms (ms)
static int time_dp_fill(gta_ctx* c, const DevState& S, const std::vector<unsigned char>& form, float& ms) {
    DpBuckets B;
    int rc = upload_dp_buckets(c, form, B);
    if (rc != GTA_OK) return rc;
    const DevChunk Q = c->q.view(), R = c->r.view();
    const DpLaunch L{DPM_GRAND_L, 1, 0.0f, 0, 0.0f, 0};
    hipStream_t st = c->stream;
    HIPCHK(c, launch_fill_i32(B.failflag, 0, 1, st));
    HIPCHK(c, launch_dp(Q, R, S, L, B, WorkQ{}, 0, st));
    HIPCHK(c, hipEventRecord(c->ev0, st));
    for (int k = 0; k < 2; k++) HIPCHK(c, launch_dp(Q, R, S, L, B, WorkQ{}, 0, st));
    HIPCHK(c, hipEventRecord(c->ev1, st));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    ms *= 0.5f;
    int dpfail = 0;
    HIPCHK(c, hipMemcpy(&dpfail, B.failflag, 4, hipMemcpyDeviceToHost));
    if (dpfail) return fail(c, GTA_ERR_HIP, "DP autotuning: a row wait exceeded its spin limit");
    return GTA_OK;
}

static int tune_dp_forms(gta_ctx* c, const DevState& S, std::vector<unsigned char>& form) {
    const int nq = c->q.nstrs, nr = c->r.nstrs;
    struct Bucket {
        std::vector<int> pairs;
        long long cells = 0, maxcells = 0;
        int minstrips = 1 << 30;
    };
    std::map<std::pair<int, int>, Bucket> buckets;
    for (int qi = 0; qi < nq; qi++) {
        const int Lq = c->q.len[qi], ns = (Lq + 63) / 64;
        for (int ri = 0; ri < nr; ri++) {
            const int Lr = c->r.len[ri];
            const long long cells = (long long)Lq * Lr;
            if (ns < 4 || cells < (1ll << 18)) continue;
            Bucket& b = buckets[{ilog2ll(ns), ilog2ll(Lr)}];
            b.pairs.push_back(qi * nr + ri);
            b.cells += cells;
            b.maxcells = std::max(b.maxcells, cells);
            b.minstrips = std::min(b.minstrips, ns);
        }
    }
    std::vector<std::pair<std::string, Bucket*>> order;
    for (auto& kv : buckets) {
        const std::string key = std::to_string(kv.first.first) + " " + std::to_string(kv.first.second) + " " +
                                std::to_string(ilog2ll((long long)kv.second.pairs.size()));
        order.push_back({key, &kv.second});
    }
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.second->cells > b.second->cells; });
    bool prepared = false;
    for (auto& kb : order) {
        Bucket& b = *kb.second;
        auto it = c->tune.find(kb.first);
        if (it != c->tune.end()) {
            for (int pidx : b.pairs) form[pidx] = (unsigned char)it->second;
            continue;
        }
        const auto t0 = std::chrono::steady_clock::now();
        if (!prepared) {
            // the fill's cost does not depend on the transform: identity for every pair
            const DevChunk Q = c->q.view(), R = c->r.view();
            HIPCHK(c, launch_init(Q, R, S, c->stream));
            std::vector<float> id((size_t)nq * nr * 12, 0.f);
            for (size_t p = 0; p < (size_t)nq * nr; p++) id[p * 12] = id[p * 12 + 4] = id[p * 12 + 8] = 1.f;
            HIPCHK(c, hipMemcpyAsync(S.tfm_grand, id.data(), id.size() * 4, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            prepared = true;
        }
        float tms[DPF_N];
        int best = -1;
        for (int f = 0; f < DPF_N; f++) {
            tms[f] = -1.f;
            if (f == DPF_ONE && b.maxcells > (1ll << 27)) continue;        // one wave would take seconds
            if (f >= DPF_W4 && f <= DPF_W16 && (4 << (f - DPF_W4)) > b.minstrips) continue;
            for (int pidx : b.pairs) form[pidx] = (unsigned char)f;
            int rc = time_dp_fill(c, S, form, tms[f]);
            if (rc != GTA_OK) return rc;
            if (best < 0 || tms[f] < tms[best]) best = f;
        }
        for (int pidx : b.pairs) form[pidx] = (unsigned char)best;
        c->tune[kb.first] = best;
        if (!c->tune_path.empty()) {
            if (FILE* fp = fopen(c->tune_path.c_str(), "a")) {
                fprintf(fp, "%s %s %d", c->tune_tag.c_str(), kb.first.c_str(), best);
                for (int f = 0; f < DPF_N; f++) fprintf(fp, " %.3f", tms[f]);
                fprintf(fp, "\n");
                fclose(fp);
            }
        }
        c->tune_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return GTA_OK;
}

static int build_dp_buckets(gta_ctx* c, const DevState& S) {
    const size_t np = (size_t)c->q.nstrs * c->r.nstrs;
    c->dpform.assign(np, DPF_ONE);
    default_dp_forms(c, c->dpform);
    if (c->autotune) {
        int rc = tune_dp_forms(c, S, c->dpform);
        if (rc != GTA_OK) return rc;
    }
    return upload_dp_buckets(c, c->dpform, c->dpb);
}

// device buffers of the pair state for the resident chunks
static int prepare_state(gta_ctx* c, DevState& S, size_t& alnsize) {
    const int nq = c->q.nstrs, nr = c->r.nstrs;
    const size_t np = (size_t)nq * nr;
    hipStream_t st = c->stream;
    // pair state
    HIPCHK(c, c->s_flags.ensure(np * 4));
    HIPCHK(c, c->s_grand.ensure(np * 4));
    HIPCHK(c, c->s_best0.ensure(np * 4));
    HIPCHK(c, c->s_prev.ensure(np * 4));
    HIPCHK(c, c->s_naln.ensure(np * 4));
    HIPCHK(c, c->s_qpos.ensure(np * 4));
    HIPCHK(c, c->s_rpos.ensure(np * 4));
    HIPCHK(c, c->s_fndx.ensure(np * 4));
    HIPCHK(c, c->s_fpos.ensure(np * 4));
    HIPCHK(c, c->s_tibest.ensure(np * 12 * 4));
    HIPCHK(c, c->s_tgrand.ensure(np * 12 * 4));
    HIPCHK(c, c->s_out.ensure(np * 15 * 4));
    const size_t alnplane = (size_t)nq * c->r.nposs;
    HIPCHK(c, c->s_aln.ensure(alnplane * 9 * 4));
    HIPCHK(c, c->s_cells.ensure(16));
    HIPCHK(c, c->s_algb.ensure(KC_COUNT * 16));
    // backtrace: per pair ceil(Lq/64) strips x ceil((Lr+63)/16) blocks x 64 words
    std::vector<unsigned long long> off(np);
    unsigned long long tot = 0;
    for (int qi = 0; qi < nq; qi++) {
        const unsigned long long ns = (unsigned long long)((c->q.len[qi] + 63) / 64);
        for (int ri = 0; ri < nr; ri++) {
            off[(size_t)qi * nr + ri] = tot;
            tot += ns * (unsigned long long)((c->r.len[ri] + 63 + 15) / 16) * 64ull;
        }
    }
    HIPCHK(c, c->s_btck.ensure(tot * 4));
    HIPCHK(c, c->s_btckoff.ensure(np * 8));
    HIPCHK(c, hipMemcpyAsync(c->s_btckoff.p, off.data(), np * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, c->s_rowbuf.ensure(alnplane * 4));
    alnsize = gta_alignment_section_size(nq, c->q.nposs, nr, c->r.nposs);
    HIPCHK(c, c->s_alns.ensure((alnsize + 3) & ~(size_t)3));

    S.nq = nq; S.nr = nr;
    S.flags = c->s_flags.as<int>(); S.grand = c->s_grand.as<float>(); S.best0 = c->s_best0.as<float>();
    S.prevbest0 = c->s_prev.as<float>(); S.nalnposs = c->s_naln.as<int>();
    S.qrypos = c->s_qpos.as<int>(); S.rfnpos = c->s_rpos.as<int>();
    S.sfragndx = c->s_fndx.as<int>(); S.sfragpos = c->s_fpos.as<int>();
    S.tfm_ibest = c->s_tibest.as<float>(); S.tfm_grand = c->s_tgrand.as<float>();
    S.out = c->s_out.as<float>();
    S.aln = c->s_aln.as<float>(); S.alnplane = alnplane;
    S.btck = c->s_btck.as<unsigned>(); S.btckoff = c->s_btckoff.as<unsigned long long>();
    S.rowbuf = c->s_rowbuf.as<float>();
    S.cells = c->s_cells.as<unsigned long long>();
    S.algbytes = c->s_algb.as<unsigned long long>();
    S.algflops = S.algbytes + KC_COUNT;
    // DP launch forms (one wave, W waves or the whole device per pair)
    {
        int rc = build_dp_buckets(c, S);
        if (rc != GTA_OK) return rc;
    }
    // refinement split
    {
        std::vector<int> small, big;
        for (int qi = 0; qi < nq; qi++)
            for (int ri = 0; ri < nr; ri++) {
                const int pidx = qi * nr + ri;
                if (std::min(c->q.len[qi], c->r.len[ri]) >= c->split_min_len) big.push_back(pidx);
                else small.push_back(pidx);
            }
        RefSplit Sp{};
        Sp.seeds_per_part = c->split_seeds;
        Sp.slots_per_part = c->split_slots;
        if (!big.empty()) {
            const Geom G = make_geom(c);
            const int maxparts = std::max({(G.nseeds + c->split_seeds - 1) / c->split_seeds,
                                           (G.nlocsteps + c->split_slots - 1) / c->split_slots,
                                           (G.maxnfragfcts * NMAXSUBFRAGS + c->split_slots - 1) / c->split_slots,
                                           1});
            HIPCHK(c, c->s_rsmall.ensure(std::max<size_t>(small.size(), 1) * 4));
            HIPCHK(c, c->s_rbig.ensure(big.size() * 4));
            HIPCHK(c, c->s_parts.ensure(big.size() * (size_t)maxparts * part_record_bytes()));
            if (!small.empty())
                HIPCHK(c, hipMemcpy(c->s_rsmall.p, small.data(), small.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(c->s_rbig.p, big.data(), big.size() * 4, hipMemcpyHostToDevice));
            Sp.small = c->s_rsmall.as<int>();
            Sp.nsmall = (int)small.size();
            Sp.big = c->s_rbig.as<int>();
            Sp.nbig = (int)big.size();
            Sp.partbuf = c->s_parts.p;
        }
        c->rsp = Sp;
    }
    // per-pair bucket marks and the live-pair lists (both forms list the pairs
    // left after each filter; the batched form drains the lists with
    // persistent grids, the plain form launches one workgroup per listed pair)
    {
        c->bigmask = nullptr;
        if (c->dpb.any_big() || c->rsp.nbig > 0) {
            std::vector<unsigned char> m(np, 0);
            for (size_t pidx = 0; pidx < np; pidx++)
                if (c->dpform[pidx] != DPF_ONE) m[pidx] |= 1;
            if (c->rsp.nbig > 0) {
                std::vector<int> big(c->rsp.nbig);
                HIPCHK(c, hipMemcpy(big.data(), c->rsp.big, big.size() * 4, hipMemcpyDeviceToHost));
                for (int pidx : big) m[pidx] |= 2;
            }
            HIPCHK(c, c->s_bigmask.ensure(np));
            HIPCHK(c, hipMemcpy(c->s_bigmask.p, m.data(), np, hipMemcpyHostToDevice));
            c->bigmask = c->s_bigmask.as<unsigned char>();
        }
        HIPCHK(c, c->s_lista.ensure(np * 4));
        HIPCHK(c, c->s_listb.ensure(np * 4));
        if (c->dpb.any_big()) HIPCHK(c, c->s_listc.ensure(np * 4));
        HIPCHK(c, c->s_lcnt.ensure(((np + 1023) / 1024) * 12));
        HIPCHK(c, c->s_ltot.ensure(12));
        HIPCHK(c, c->s_heads.ensure(GTA_MAX_QUEUES * 4));
    }

    return GTA_OK;
}

// Waves per workgroup of the DPRefine launches.  One-wave workgroups leave
// fewer of a CU's lanes idle while a pair's last slots run, but stage only
// alignments of up to rlds_cap32(1) positions in LDS; longer ones are read from
// L2 in every pass.  An alignment is at most min(Lq, Lr) long, so one wave is
// used when at most 15% of the chunks' pairs can have longer alignments (16 x
// 50,000 workload: 7%, DPRefine 0.85 -> 0.79 s with one wave; 1 x 500,000: 27%,
// 0.40 -> 0.50 s).
static int dprefine_waves(const gta_ctx* c) {
    const int cap = rlds_cap32(1);
    long long ql = 0, rl = 0;
    for (int v : c->q.len) ql += v > cap;
    for (int v : c->r.len) rl += v > cap;
    const double np = (double)c->q.len.size() * (double)c->r.len.size();
    return (np > 0 && (double)ql * (double)rl <= 0.15 * np) ? 1 : 2;
}

static int run_stage(gta_ctx* c, double* seconds, double* cells) {
    if (!c) return fail(nullptr, GTA_ERR_ARG, "null context");
    if (!c->have_q || !c->have_r) return fail(c, GTA_ERR_STATE, "queries and references must be set first");
    HIPCHK(c, hipSetDevice(c->device));
    const int nq = c->q.nstrs, nr = c->r.nstrs;
    const size_t np = (size_t)nq * nr;
    const int prec = c->opts.precision;
    const int dinv = c->opts.symmetric;

    DevState S;
    size_t alnsize = 0;
    {
        int rc0 = prepare_state(c, S, alnsize);
        if (rc0 != GTA_OK) return rc0;
    }
    const DevChunk Q = c->q.view(), R = c->r.view();
    const Geom G = make_geom(c);
    const gta_options& o = c->opts;
    hipStream_t st = c->stream;

    c->evrec.clear();
    size_t evnext = 0;
    // brackets one launch with events when profiling (kernel class kc)
    auto ev_begin = [&](int kc) -> hipError_t {
        if (!c->profiling) return hipSuccess;
        while (c->evpool.size() < evnext + 2) {
            hipEvent_t e;
            hipError_t er = hipEventCreate(&e);
            if (er != hipSuccess) return er;
            c->evpool.push_back(e);
        }
        c->evrec.push_back({kc, (int)evnext});
        hipError_t er = hipEventRecord(c->evpool[evnext], st);
        evnext += 2;
        return er;
    };
    auto ev_end = [&]() -> hipError_t {
        if (!c->profiling) return hipSuccess;
        return hipEventRecord(c->evpool[c->evrec.back().second + 1], st);
    };
#define GTA_TIMED(kc, expr)                  \
    do {                                     \
        HIPCHK(c, ev_begin(kc));             \
        HIPCHK(c, (expr));                   \
        HIPCHK(c, ev_end());                 \
    } while (0)
    int maxaln = std::min(c->q.len.front(), c->r.len.front());
    RefineLaunch RL{dinv, 0, 0, 0, 0.f, std::min(maxaln, 1536), dprefine_waves(c)};
    // After each filter (the stage-1 prescreen, the -s threshold) the live
    // pairs are listed on the device.  The batched form's launches that follow
    // are persistent grids draining those lists (no host round trip: the
    // sequence replays as a HIP graph); the plain form reads the list lengths
    // back and launches one workgroup per listed pair.
    const bool batched = o.batched != 0;
    bool listed = false;
    int nqueue = 0;
    int* lista = c->s_lista.as<int>();
    int* listb = c->s_listb.as<int>();
    // every live pair, when the one-wave DP list leaves some out
    int* listc = c->dpb.any_big() ? c->s_listc.as<int>() : nullptr;
    int* ltot = c->s_ltot.as<int>();
    int* heads = batched ? c->s_heads.as<int>() : nullptr;
    const int pgrid_dp = c->ncu * 32, pgrid_rf = c->ncu * REFINE_WGS_PER_CU;
    DpBuckets dpb = c->dpb;         // plain form: one-wave list narrowed to live pairs
    RefSplit rsp = c->rsp;
    auto queue = [&](const int* items, const int* count) -> WorkQ {
        if (!batched || !listed || nqueue >= GTA_MAX_QUEUES) return WorkQ{};
        return WorkQ{items, count, heads + nqueue++};
    };
    auto relist = [&]() -> hipError_t {
        listed = true;
        hipError_t e = launch_live_lists(S, c->bigmask, c->s_lcnt.as<int>(), ltot, lista, listb, listc, st);
        if (e != hipSuccess || batched) return e;
        int n[3];
        if ((e = hipMemcpyAsync(n, ltot, sizeof n, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        dpb.small = lista; dpb.nsmall = n[0];
        rsp.small = listb; rsp.nsmall = n[1];
        return hipSuccess;
    };
    auto dp = [&](int mode, int gap0, float gapopen, int usess, int dpscan) -> hipError_t {
        DpLaunch L{mode, gap0, gapopen, usess, usess ? 0.5f : 0.0f, dpscan};
        return launch_dp(Q, R, S, L, dpb, queue(lista, ltot), pgrid_dp, st);
    };
    auto refine = [&](int mode, int epi, int dpi, int maxdpi, float prescore) -> hipError_t {
        RefineLaunch L = RL;
        L.epi = epi; L.dpi = dpi; L.maxndpiters = maxdpi; L.prescore = prescore;
        // a persistent grid keeps 16 waves per CU whatever the workgroup size
        const int pg = (mode == RM_DPREFINE_L && L.waves == 1) ? c->ncu * 16 : pgrid_rf;
        return launch_refine_dp(prec, mode, Q, R, S, G, L, rsp, queue(listb, ltot + 1), pg, st);
    };
    // DPRefine<GAP0,PRESCREEN> (MpStage1.cpp:455-510)
    auto dprefine = [&](int gap0, float prescore, int maxdpi) -> int {
        for (int dpi = 0; dpi < maxdpi; dpi++) {
            GTA_TIMED(KC_DP, dp(DPM_IBEST_L, gap0, gap0 ? 0.0f : -0.6f, 0, 0));
            GTA_TIMED(KC_REFINE_DP, refine(RM_DPREFINE_L, 1, dpi, maxdpi, prescore));
        }
        return GTA_OK;
    };

    // the batched form replays the sequence below as a HIP graph once captured
    // (not while profiling: the per-launch events are host-timed)
    const bool graph = batched && !c->profiling;
    auto enqueue = [&]() -> int {
        // cleared by kernels (no memset nodes in a captured graph)
        HIPCHK(c, launch_fill_i32((int*)c->s_cells.p, 0, 4, st));
        HIPCHK(c, launch_fill_i32((int*)c->s_algb.p, 0, KC_COUNT * 4, st));
        HIPCHK(c, launch_fill_i32((int*)c->s_alns.p, 0, (alnsize + 3) / 4, st));
        HIPCHK(c, launch_fill_i32(c->dpb.failflag, 0, 1, st));
        if (batched) HIPCHK(c, launch_fill_i32(heads, 0, GTA_MAX_QUEUES, st));
        GTA_TIMED(KC_SECSTR, launch_secstr(Q, st));
        GTA_TIMED(KC_SECSTR, launch_secstr(R, st));
        GTA_TIMED(KC_INIT, launch_init(Q, R, S, st));
        // stage 1 (MpStage1.h:78-85)
        GTA_TIMED(KC_FINDFRAG, launch_findfrag(Q, R, S, G, dinv, c->rsp, st));
        GTA_TIMED(KC_REFINE_GAPLESS, launch_refine_gapless(Q, R, S, G, dinv, c->rsp, st));
        int rc;
        if ((rc = dprefine(1, o.prescore, 2)) != GTA_OK) return rc;
        // references kept for a block of queries continue with all its queries
        if (0.0f < o.prescore) GTA_TIMED(KC_MISC, launch_keep_refs(Q, R, S, o.qgroup, st));
        HIPCHK(c, relist());
        if (o.addsearchbyss) {
            // superposition search from the alignment of secondary structure
            // and residue similarity (MpStageSSRR.h:62-76, MpBatch.cpp:421-436):
            // gap open -1, SS match weight 1, residue score weight 0.2
            const DpLaunch L{DPM_IBEST_L, 0, -1.0f, 2, 1.0f, 0, 0.2f};
            GTA_TIMED(KC_DP, launch_dp(Q, R, S, L, dpb, queue(lista, ltot), pgrid_dp, st));
            GTA_TIMED(KC_REFINE_DP, refine(RM_DPREFINE_L, 0, 0, 0, 0.f));
            if ((rc = dprefine(0, 0.f, 2)) != GTA_OK) return rc;
            if ((rc = dprefine(1, 0.f, 2)) != GTA_OK) return rc;
        }
        // stage 2, pass 1: SS-weighted DP with gap cost (MpStage2.h:60-80)
        GTA_TIMED(KC_DP, dp(DPM_GRAND_L, 0, -1.0f, 1, 1));
        GTA_TIMED(KC_REFINE_DP, refine(RM_DPREFINE_L, 0, 0, 0, 0.f));
        if ((rc = dprefine(0, 0.f, 2)) != GTA_OK) return rc;
        if ((rc = dprefine(1, 0.f, 2)) != GTA_OK) return rc;
        // stage 2, pass 2: no gap cost, convergence iterations
        GTA_TIMED(KC_DP, dp(DPM_GRAND_L, 1, 0.0f, 0, 0));
        GTA_TIMED(KC_REFINE_DP, refine(RM_DPREFINE_L, 0, 0, 0, 0.f));
        if ((rc = dprefine(0, 0.f, 2)) != GTA_OK) return rc;
        if ((rc = dprefine(1, 0.f, o.convergence)) != GTA_OK) return rc;
        if (0.0f < o.scorethld) {
            GTA_TIMED(KC_MISC, launch_setlow(Q, R, S, o.scorethld, st));
            HIPCHK(c, relist());
        }
        // final stage (MpStageFin.h:66-90)
        GTA_TIMED(KC_DP, dp(DPM_GRAND_L, 1, 0.0f, 0, 0));
        GTA_TIMED(KC_REFINE_FIN, refine(RM_FIN1_L, 0, 0, 0, 0.f));
        // MpStageFin::Refine (MpStageFin.h:337-392)
        if (!(o.depth == GTA_DEPTH_SHALLOW || o.refinement == GTA_REF_COARSE))
            GTA_TIMED(KC_REFINE_FIN, refine(o.refinement == GTA_REF_FULLA ? RM_FIN2F_L : RM_FIN2_L, 0, 0, 0, 0.f));
        GTA_TIMED(KC_DP, dp(DPM_CONSTRAINED_L, 1, 0.0f, 0, 0));
        GTA_TIMED(KC_PROD_ALN, launch_prod_aln(Q, R, S, c->s_alns.as<char>(), o.nodeletions,
                                               listc ? queue(listc, ltot + 2) : queue(lista, ltot),
                                               pgrid_dp, st));
        GTA_TIMED(KC_REFINE_FIN, refine(RM_PROD1_L, 0, 0, 0, 0.f));
        if (prec) GTA_TIMED(KC_REFINE_FIN, launch_rmsd_fp64(Q, R, S, st));
        // MpStageFin::ProduceOutputScores (MpStageFin.h:398-460)
        if (o.refinement == GTA_REF_ONE || o.refinement == GTA_REF_COARSE)
            GTA_TIMED(KC_REFINE_FIN, refine(RM_PROD2_L, 0, 0, 0, 0.f));
        else if (o.refinement == GTA_REF_LOG)
            GTA_TIMED(KC_REFINE_FIN, refine(RM_PROD2L_L, 0, 0, 0, 0.f));
        else
            GTA_TIMED(KC_REFINE_FIN, refine(RM_PROD2F_L, 0, 0, 0, 0.f));
        if (o.referenced) GTA_TIMED(KC_MISC, launch_revert(S, st));
        return GTA_OK;
    };
#undef GTA_TIMED
    HIPCHK(c, hipEventRecord(c->ev0, st));
    if (graph && c->gvalid) {
        HIPCHK(c, hipGraphLaunch(c->gexec, st));
    } else {
        // ends a capture that an error return leaves open
        struct CaptureGuard {
            hipStream_t st;
            bool on = false;
            ~CaptureGuard() {
                if (on) {
                    hipGraph_t g = nullptr;
                    if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
                }
            }
        } cap{st};
        if (graph) {
            HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            cap.on = true;
        }
        const int erc = enqueue();
        if (erc != GTA_OK) return erc;
        if (graph) {
            hipGraph_t g = nullptr;
            cap.on = false;
            HIPCHK(c, hipStreamEndCapture(st, &g));
            if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
            hipError_t ge = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(c, ge);
            c->gvalid = true;
            HIPCHK(c, hipGraphLaunch(c->gexec, st));
        }
    }
    HIPCHK(c, hipEventRecord(c->ev1, st));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    unsigned long long ncells[2] = {0, 0};
    HIPCHK(c, hipMemcpy(ncells, c->s_cells.p, 16, hipMemcpyDeviceToHost));
    c->cells_useful = (double)ncells[0];
    c->cells_launched = (double)ncells[1];
    if (seconds) *seconds = ms * 1e-3;
    if (cells) *cells = (double)ncells[0];
    for (int k = 0; k < KC_COUNT; k++) { c->kc_launches[k] = 0; c->kc_seconds[k] = 0; }
    if (c->profiling) {
        for (const auto& e : c->evrec) {
            float t = 0.f;
            HIPCHK(c, hipEventElapsedTime(&t, c->evpool[e.second], c->evpool[e.second + 1]));
            c->kc_launches[e.first]++;
            c->kc_seconds[e.first] += t * 1e-3;
        }
    }
    int dpfail = 0;
    HIPCHK(c, hipMemcpy(&dpfail, c->s_fail.p, 4, hipMemcpyDeviceToHost));
    if (dpfail) return fail(c, GTA_ERR_HIP, "multi-wave or grid DP: a row wait exceeded its spin limit");
    unsigned long long ab[2 * KC_COUNT];
    HIPCHK(c, hipMemcpy(ab, c->s_algb.p, sizeof(ab), hipMemcpyDeviceToHost));
    for (int k = 0; k < KC_COUNT; k++) {
        c->kc_bytes[k] = (double)ab[k];
        c->kc_flops[k] = (double)ab[KC_COUNT + k];
    }
    return GTA_OK;
}

int gta_set_profiling(gta_ctx* c, int on) {
    if (!c) return fail(nullptr, GTA_ERR_ARG, "null context");
    c->profiling = on ? 1 : 0;
    return GTA_OK;
}

int gta_dp_cell_counts(const gta_ctx* c, double* useful, double* launched) {
    if (!c || !useful || !launched) return GTA_ERR_ARG;
    *useful = c->cells_useful;
    *launched = c->cells_launched;
    return GTA_OK;
}

int gta_kernel_stats(const gta_ctx* c, gta_kernel_stat* out, int n) {
    static const char* names[KC_COUNT] = {"secstr", "init", "findfrag", "refine_gapless", "dp",
                                          "refine_dp", "refine_final", "prod_aln", "misc"};
    if (!c) return GTA_ERR_ARG;
    for (int k = 0; k < KC_COUNT && k < n; k++) {
        out[k].name = names[k];
        out[k].launches = c->kc_launches[k];
        out[k].seconds = c->kc_seconds[k];
        out[k].alg_bytes = c->kc_bytes[k];
        out[k].alg_flops = c->kc_flops[k];
    }
    return KC_COUNT;
}

int gta_test_dp(gta_ctx* c, const float* tfms, int gap0, float gapopen, int usess, int dpscan,
                int* nmatched, int* qpos, int* rpos) {
    if (!c || !tfms || !nmatched || !qpos || !rpos) return fail(c, GTA_ERR_ARG, "null argument");
    if (!c->have_q || !c->have_r) return fail(c, GTA_ERR_STATE, "queries and references must be set first");
    HIPCHK(c, hipSetDevice(c->device));
    DevState S;
    size_t alnsize = 0;
    int rc = prepare_state(c, S, alnsize);
    if (rc != GTA_OK) return rc;
    const DevChunk Q = c->q.view(), R = c->r.view();
    const size_t np = (size_t)c->q.nstrs * c->r.nstrs;
    hipStream_t st = c->stream;
    HIPCHK(c, launch_secstr(Q, st));
    HIPCHK(c, launch_secstr(R, st));
    HIPCHK(c, launch_init(Q, R, S, st));
    HIPCHK(c, hipMemcpyAsync(S.tfm_grand, tfms, np * 12 * 4, hipMemcpyHostToDevice, st));
    if (usess < 0 || usess > 2) return fail(c, GTA_ERR_ARG, "usess must be 0, 1 or 2");
    // kind 2: the SS/residue fill with the weights of the --add-search-by-ss search
    const DpLaunch L{DPM_GRAND_L, gap0, gapopen, usess, usess == 2 ? 1.0f : usess ? 0.5f : 0.0f, dpscan,
                     usess == 2 ? 0.2f : 0.0f};
    HIPCHK(c, launch_fill_i32(c->dpb.failflag, 0, 1, st));
    HIPCHK(c, launch_dp(Q, R, S, L, c->dpb, WorkQ{}, 0, st));
    HIPCHK(c, hipStreamSynchronize(st));
    int dpfail = 0;
    HIPCHK(c, hipMemcpy(&dpfail, c->dpb.failflag, 4, hipMemcpyDeviceToHost));
    if (dpfail) return fail(c, GTA_ERR_HIP, "multi-wave or grid DP: a row wait exceeded its spin limit");
    std::vector<float> pl(S.alnplane);
    HIPCHK(c, hipMemcpy(nmatched, S.nalnposs, np * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(pl.data(), S.aln + 7 * S.alnplane, S.alnplane * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < S.alnplane; i++) qpos[i] = (int)pl[i];
    HIPCHK(c, hipMemcpy(pl.data(), S.aln + 8 * S.alnplane, S.alnplane * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < S.alnplane; i++) rpos[i] = (int)pl[i];
    return GTA_OK;
}

int gta_test_dp_score(gta_ctx* c, const float* tfms, const int* slotflags, int nslots, int gap0, float gapopen,
                      float* scores) {
    if (!c || !tfms || !scores || nslots < 1) return fail(c, GTA_ERR_ARG, "invalid argument");
    if (!c->have_q || !c->have_r) return fail(c, GTA_ERR_STATE, "queries and references must be set first");
    HIPCHK(c, hipSetDevice(c->device));
    DevState S;
    size_t alnsize = 0;
    int rc = prepare_state(c, S, alnsize);
    if (rc != GTA_OK) return rc;
    const DevChunk Q = c->q.view(), R = c->r.view();
    const size_t np = (size_t)c->q.nstrs * c->r.nstrs, n = np * (size_t)nslots;
    hipStream_t st = c->stream;
    HIPCHK(c, c->s_sctfm.ensure(n * 12 * 4));
    HIPCHK(c, c->s_scout.ensure(n * 4));
    HIPCHK(c, hipMemcpyAsync(c->s_sctfm.p, tfms, n * 12 * 4, hipMemcpyHostToDevice, st));
    const int* dflags = nullptr;
    if (slotflags) {
        HIPCHK(c, c->s_scflag.ensure(n * 4));
        HIPCHK(c, hipMemcpyAsync(c->s_scflag.p, slotflags, n * 4, hipMemcpyHostToDevice, st));
        dflags = c->s_scflag.as<int>();
    }
    HIPCHK(c, launch_init(Q, R, S, st));
    HIPCHK(c, launch_dp_score(Q, R, S, gap0, gapopen, c->s_sctfm.as<float>(), dflags, nslots,
                              c->s_scout.as<float>(), st));
    HIPCHK(c, hipStreamSynchronize(st));
    HIPCHK(c, hipMemcpy(scores, c->s_scout.p, n * 4, hipMemcpyDeviceToHost));
    return GTA_OK;
}

int gta_test_set_dp_buckets(gta_ctx* c, long long min_cells, int min_strips) {
    if (!c || min_cells < 0 || min_strips < 1) return fail(c, GTA_ERR_ARG, "invalid DP bucket thresholds");
    c->mw_min_cells = min_cells;
    c->mw_min_strips = min_strips;
    c->gvalid = false;   // the captured graph holds the old launch forms and buffers
    return GTA_OK;
}

int gta_set_autotune(gta_ctx* c, int on, const char* cache_path) {
    if (!c) return fail(nullptr, GTA_ERR_ARG, "null context");
    c->autotune = on ? 1 : 0;
    c->gvalid = false;
    c->tune.clear();
    c->tune_path = cache_path ? cache_path : "";
    if (c->tune_tag.empty()) {
        hipDeviceProp_t prop;
        HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
        c->tune_tag = std::string(prop.gcnArchName) + "/" + std::to_string(c->ncu) + "cu";
        for (char& ch : c->tune_tag)
            if (ch == ' ') ch = '_';
    }
    if (!c->tune_path.empty()) {
        if (FILE* fp = fopen(c->tune_path.c_str(), "r")) {
            char tag[256];
            int sc, lc, cc, f;
            char line[1024];
            while (fgets(line, sizeof line, fp)) {
                if (sscanf(line, "%255s %d %d %d %d", tag, &sc, &lc, &cc, &f) == 5 && c->tune_tag == tag &&
                    f >= 0 && f < DPF_N)
                    c->tune[std::to_string(sc) + " " + std::to_string(lc) + " " + std::to_string(cc)] = f;
            }
            fclose(fp);
        }
    }
    return GTA_OK;
}

int gta_test_dp_forms(const gta_ctx* c, int* counts, double* tune_seconds) {
    if (!c || !counts) return GTA_ERR_ARG;
    for (int f = 0; f < DPF_N; f++) counts[f] = 0;
    for (unsigned char f : c->dpform) counts[f]++;
    if (tune_seconds) *tune_seconds = c->tune_seconds;
    return GTA_OK;
}

int gta_test_set_dp_grid(gta_ctx* c, int mode) {
    if (!c || mode < 0 || mode > 2) return fail(c, GTA_ERR_ARG, "invalid DP grid mode");
    c->dp_grid = mode;
    c->gvalid = false;   // the captured graph holds the old launch forms and buffers
    return GTA_OK;
}

int gta_test_pair_score(gta_ctx* c, const float* d02, const float* d2, int n, float* out) {
    if (!c || !d02 || !d2 || !out || n < 0) return fail(c, GTA_ERR_ARG, "invalid argument");
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf b;
    const size_t sz = (size_t)n * sizeof(float);
    hipError_t e = b.ensure(3 * sz);
    if (e != hipSuccess) return fail(c, GTA_ERR_HIP, hipGetErrorString(e));
    float* a = b.as<float>();
    int rc = GTA_OK;
    if ((e = hipMemcpy(a, d02, sz, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(a + n, d2, sz, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = launch_pair_score(a, a + n, n, a + 2 * (size_t)n, c->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess ||
        (e = hipMemcpy(out, a + 2 * (size_t)n, sz, hipMemcpyDeviceToHost)) != hipSuccess)
        rc = fail(c, GTA_ERR_HIP, hipGetErrorString(e));
    b.release();
    return rc;
}

int gta_test_valu_calibration(gta_ctx* c, int blocks, int iters, int kind, int lanes, double* seconds) {
    if (!c || blocks < 1 || iters < 1 || kind < 0 || kind > 7 || lanes < 1 || lanes > 64 || !seconds)
        return fail(c, GTA_ERR_ARG, "invalid argument");
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf b;
    hipError_t e = b.ensure((size_t)blocks * 256 * sizeof(float));
    if (e != hipSuccess) return fail(c, GTA_ERR_HIP, hipGetErrorString(e));
    float ms = 0.f;
    int rc = GTA_OK;
    if ((e = launch_valu_calib(b.as<float>(), blocks, 1, kind, lanes, c->stream)) != hipSuccess ||
        (e = hipEventRecord(c->ev0, c->stream)) != hipSuccess ||
        (e = launch_valu_calib(b.as<float>(), blocks, iters, kind, lanes, c->stream)) != hipSuccess ||
        (e = hipEventRecord(c->ev1, c->stream)) != hipSuccess ||
        (e = hipEventSynchronize(c->ev1)) != hipSuccess ||
        (e = hipEventElapsedTime(&ms, c->ev0, c->ev1)) != hipSuccess)
        rc = fail(c, GTA_ERR_HIP, hipGetErrorString(e));
    *seconds = ms * 1e-3;
    (void)hipStreamSynchronize(c->stream);
    b.release();
    return rc;
}

int gta_test_set_split(gta_ctx* c, int min_len, int seeds_per_part, int slots_per_part) {
    if (!c || min_len < 1 || seeds_per_part < 1 || slots_per_part < 1)
        return fail(c, GTA_ERR_ARG, "invalid split parameters");
    c->split_min_len = min_len;
    c->split_seeds = seeds_per_part;
    c->split_slots = slots_per_part;
    c->gvalid = false;   // the captured graph holds the old split lists and buffers
    return GTA_OK;
}

int gta_process_resident(gta_ctx* c, double* seconds, double* dp_cells) {
    return run_stage(c, seconds, dp_cells);
}

int gta_process(gta_ctx* c, gta_results* out) {
    if (!out) return fail(c, GTA_ERR_ARG, "null results pointer");
    double secs = 0, cells = 0;
    int rc = run_stage(c, &secs, &cells);
    if (rc != GTA_OK) return rc;
    const int nq = c->q.nstrs, nr = c->r.nstrs;
    const size_t np = (size_t)nq * nr;
    const size_t alnsize = gta_alignment_section_size(nq, c->q.nposs, nr, c->r.nposs);
    c->h_out.resize(np * 15);
    c->h_tfm.resize(np * 12);
    c->h_alns.resize(alnsize);
    HIPCHK(c, hipMemcpy(c->h_out.data(), c->s_out.p, np * 15 * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(c->h_tfm.data(), c->s_tgrand.p, np * 12 * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(c->h_alns.data(), c->s_alns.p, alnsize, hipMemcpyDeviceToHost));
    out->nqrys = nq;
    out->nrefs = nr;
    out->alndata = c->h_out.data();
    out->tfms = c->h_tfm.data();
    out->alns = c->h_alns.data();
    out->alnsize = alnsize;
    out->seconds = secs;
    out->dp_cells = cells;
    return GTA_OK;
}

int gta_process_noalloc(gta_ctx* c, const gta_chunk_view* queries, const gta_chunk_view* refs, float* alndata,
                        float* tfms, char* alns, size_t alns_bytes) {
    if (!c) return fail(nullptr, GTA_ERR_ARG, "null context");
    if (!queries || !refs || !alndata || !tfms || !alns) return fail(c, GTA_ERR_ARG, "null argument");
    if (!c->scratch) return fail(c, GTA_ERR_HIP, "no scratch address for this device");
    const size_t need = gta_alignment_section_size(queries->nstrs, queries->nposs, refs->nstrs, refs->nposs);
    if (alns_bytes < need) return fail(c, GTA_ERR_ARG, "alignment buffer smaller than the alignment section");
    (void)hipSetDevice(c->device);
    // the context's own buffers step aside; every buffer of the call is carved
    // from this device's scratch
    std::vector<DevBuf*> bufs = all_bufs(c);
    std::vector<DevBuf> saved;
    saved.reserve(bufs.size());
    for (DevBuf* b : bufs) {
        saved.push_back(*b);
        b->p = nullptr;
        b->bytes = 0;
    }
    const gta_options keep = c->opts;
    const int keep_prof = c->profiling, keep_tune = c->autotune;
    c->opts.batched = 0;
    c->profiling = 0;
    c->autotune = 0;
    c->gvalid = false;
    Arena arena;
    arena.base = (char*)c->scratch;
    arena.cap = GTA_NOALLOC_SCRATCH_BYTES;
    g_arena = &arena;
    int rc = gta_set_queries(c, queries);
    if (rc == GTA_OK) rc = gta_set_references(c, refs);
    double secs = 0, cells = 0;
    if (rc == GTA_OK) rc = run_stage(c, &secs, &cells);
    if (rc == GTA_OK) {
        const size_t np = (size_t)queries->nstrs * refs->nstrs;
        hipError_t e = hipMemcpy(alndata, c->s_out.p, np * 15 * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(tfms, c->s_tgrand.p, np * 12 * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(alns, c->s_alns.p, need, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, GTA_ERR_HIP, std::string("result copy: ") + hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(c->stream);
    g_arena = nullptr;
    if (arena.over)
        rc = fail(c, GTA_ERR_SIZE, "the chunks need more device scratch than GTA_NOALLOC_SCRATCH_BYTES (" +
                                       std::to_string((unsigned long long)GTA_NOALLOC_SCRATCH_BYTES) + " bytes)");
    for (size_t i = 0; i < bufs.size(); i++) *bufs[i] = saved[i];
    c->opts = keep;
    c->profiling = keep_prof;
    c->autotune = keep_tune;
    // the resident chunks were replaced: set them again before gta_process
    c->have_q = c->have_r = false;
    c->gvalid = false;
    return rc;
}

}  // extern "C"
