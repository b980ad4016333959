/*
 * gtalign_amd.h -- C ABI of the MI355X structure-alignment stage.
 *
 * This is the drop-in boundary for the GPU batch object of GTalign
 * (CuBatch, src/libmycu/cubatch/CuBatch.cuh:38-241).  A query chunk and a
 * reference chunk of C-alpha coordinates go in; per pair TM-scores, RMSD,
 * the 3x4 superposition and the residue alignment come out, in the same
 * three result sections the reference copies to the host
 * (CuBatch.cu:769-818).
 *
 * All entry points return 0 on success and a negative gta_status on error;
 * gta_last_error() then returns a message.  No C++ exception crosses this
 * boundary.  One context per host thread; a context is not re-entrant.
 */
#ifndef GTALIGN_AMD_H
#define GTALIGN_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- layout constants (reference: src/libgenp/gproc/gproc.h) ---------- */

/* transform: row-major rotation then translation (gproc.h:15-29) */
#define GTA_TFM_NFIELDS 12
/* per-pair alignment record, field order of TDP2OutputAlnData (gproc.h:207-227) */
enum gta_aln_field {
    GTA_ALN_ORGSTRNO = 0,
    GTA_ALN_STRNEWDST,
    GTA_ALN_BEGCOORDS,   /* 1-based (ref<<16 | qry) packed into the float bits */
    GTA_ALN_ENDCOORDS,
    GTA_ALN_ALNLENGTH,
    GTA_ALN_PSTVS,       /* aligned pairs closer than 5 A */
    GTA_ALN_IDNTS,
    GTA_ALN_NGAPS,
    GTA_ALN_RMSD,        /* 9999.9 when the pair was not aligned */
    GTA_ALN_SCOREQ,      /* TM-score normalised by the query length */
    GTA_ALN_SCORER,      /* TM-score normalised by the reference length */
    GTA_ALN_D0Q,
    GTA_ALN_D0R,
    GTA_ALN_2SCOREQ,
    GTA_ALN_2SCORER,
    GTA_ALN_NFIELDS      /* = 15 */
};
/* alignment rows (TDP2OutputAlignment / ...SSS, gproc.h:230-238) */
#define GTA_ALN_NROWS 5

/* ---- chunk view: the 10 SoA fields of TPM2DVectorFields ---------------
 * (reference: src/libgenp/gdats/PM2DVectorFields.h:220-258).
 * Structures are sorted by descending length (PMBatchStrData::Sort,
 * PMBatchStrData.cpp:110-160); dst[i] is the prefix offset of structure i.
 * ss is not read: the stage assigns every position's secondary structure
 * itself, as the reference does (MpSecStr::Run, MpSecStr.h:50-56).
 * Coordinates must be finite with magnitude at most GTA_MAX_COORD (the pair
 * score d0^2/(d0^2+d^2) is divided exactly only for d^2 far below the float
 * exponent limit); gta_set_* rejects others with GTA_ERR_ARG. */
#define GTA_MAX_COORD 1.0e15f
typedef struct gta_chunk_view {
    int nstrs;
    int nposs;
    const int* len;          /* pps2DLen   */
    const unsigned* dst;     /* pps2DDist  */
    const int* type;         /* pps2DType  */
    const float* x;          /* pmv2DCoords+0 */
    const float* y;          /* pmv2DCoords+1 */
    const float* z;          /* pmv2DCoords+2 */
    const int* ins_ch_ord;   /* pmv2D_Ins_Ch_Ord: >0 marks a nucleic acid */
    const int* resnum;       /* pmv2DResNumber */
    const char* rsd;         /* pmv2Drsd: residue one-letter code */
    const char* ss;          /* pmv2Dss (not read; may be NULL) */
} gta_chunk_view;

/* ---- options: the CLOptions the stage reads (CuBatch.cu:353-361) ------ */
enum gta_depth { GTA_DEPTH_DEEP = 0, GTA_DEPTH_HIGH, GTA_DEPTH_MEDIUM, GTA_DEPTH_SHALLOW };
/* internal refinement codes (CLOptions.h:122-127; the CLI stores option+1),
 * selecting the final stage's searches (MpStageFin.h:337-460):
 *   LOG    production phase 2 as a log search of positions;
 *   COARSE no final phase 2; production phase 2 at one fragment length;
 *   ONE    final and production phase 2 at one fragment length (default);
 *   FULL   production phase 2 over every fragment length;
 *   FULLA  final and production phase 2 over every fragment length.
 * The production full search (FULL, FULLA) runs only slots at alignment
 * positions >= 0; the reference also runs slots at negative positions, on
 * buffer memory past the pair's alignment, whose outcome depends on its
 * memory history. */
enum gta_refinement { GTA_REF_LOG = 0, GTA_REF_COARSE, GTA_REF_ONE, GTA_REF_FULL, GTA_REF_FULLA };
enum gta_precision {
    GTA_PREC_FP32 = 0,       /* plain form: FP32 throughout, as the reference */
    GTA_PREC_MIXED = 1       /* searches as FP32; final refinement with FP64 covariance
                                and 3x3 solve; RMSD from an FP64 eigen-solve */
};

typedef struct gta_options {
    float prescore;          /* --pre-score   (default 0.4) */
    float scorethld;         /* -s            (default 0.5) */
    int convergence;         /* --convergence (default 18)  */
    int depth;               /* gta_depth     (default MEDIUM) */
    int refinement;          /* gta_refinement (default ONE) */
    int symmetric;           /* --symmetric   (default 0) */
    int nodeletions;         /* --no-deletions (default 0) */
    int referenced;          /* --referenced  (default 0) */
    int precision;           /* gta_precision (default FP32) */
    int batched;             /* execution form: 0 plain (one workgroup per pair in every launch),
                                1 batched (after each filter the live pairs are listed on the device
                                and persistent grids drain those lists; the launch sequence is
                                captured once as a HIP graph and replayed) -- default 0 */
    int addsearchbyss;       /* --add-search-by-ss (default 0): after stage 1, an extra superposition
                                search from an alignment of secondary structure and residue similarity
                                (MpStageSSRR.h:62-76) */
    int qgroup;              /* queries per block of the reference filter after stage 1 (default 0):
                                with --pre-score > 0 the reference drops from a block the references
                                that every query of the block flags low-scoring, and the pairs of the
                                references it keeps all continue, their flags reset (CuBatch.cu:558-563,
                                MpReform.cpp:28-124).  0: the whole query chunk is one block, as one
                                CuBatch::ProcessBlock call; k: blocks of k consecutive queries, as the
                                reference's scheduler forms them (--dev-queries-per-chunk, default 2) */
} gta_options;

/* results of one gta_process call; host memory owned by the context and
 * valid until the next gta_process / gta_destroy on it */
typedef struct gta_results {
    int nqrys;
    int nrefs;
    const float* alndata;    /* [nqrys*nrefs][GTA_ALN_NFIELDS] */
    const float* tfms;       /* [nqrys*nrefs][GTA_TFM_NFIELDS] */
    const char* alns;        /* alignment section, layout of production_match2aln.cu:130-133 */
    size_t alnsize;          /* bytes in alns */
    double seconds;          /* device time of the stage (excludes H2D/D2H) */
    double dp_cells;         /* DP cells updated in fills that ran */
} gta_results;

typedef struct gta_ctx gta_ctx;

enum gta_status {
    GTA_OK = 0,
    GTA_ERR_ARG = -1,        /* invalid argument (null pointer, bad option) */
    GTA_ERR_EMPTY = -2,      /* empty chunk (CuBatch.cu:385-386) */
    GTA_ERR_TOOMANY = -3,    /* too many queries per chunk (CuBatch.cu:397-398) */
    GTA_ERR_SIZE = -4,       /* chunk exceeds the context's capacity (CuBatch.cu:414-418) */
    GTA_ERR_LENGTH = -5,     /* structure longer than 65535 residues (PMBatchStrData.h:34) */
    GTA_ERR_HIP = -6,        /* HIP runtime error */
    GTA_ERR_STATE = -7       /* call order violated (e.g. process before set_*) */
};

/* fills opts with the reference defaults (CLOptions.cpp:35-100) */
void gta_default_options(gta_options* opts);

/* Replaces the CuBatch constructor (CuBatch.cu:54-150): binds the device,
 * creates the stream and sizes device memory for chunks of up to
 * max_qry_poss/max_ref_poss residues and max_qrys/max_refs structures. */
int gta_create(int device, const gta_options* opts,
               int max_qrys, int max_qry_poss, int max_refs, int max_ref_poss,
               gta_ctx** out);
/* Replaces ~CuBatch. Safe on NULL. */
void gta_destroy(gta_ctx* ctx);

/* Replace CuBatch::TransferQueryPMDataToDevice / TransferCPMDataToDevice
 * (CuBatch.cuh:54-72): copy one chunk's fields to HBM. */
int gta_set_queries(gta_ctx* ctx, const gta_chunk_view* qrys);
int gta_set_references(gta_ctx* ctx, const gta_chunk_view* refs);

/* Replaces CuBatch::ProcessBlock (CuBatch.cuh:74-85, CuBatch.cu:339-717):
 * runs stage 1, stage 2 (two passes) and the final stage for every
 * query x reference pair of the resident chunks and copies the three
 * result sections to host memory owned by ctx. */
int gta_process(gta_ctx* ctx, gta_results* out);

/* ---- stage without device allocation --------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory during a call (serving loops, harnesses that forbid hipMalloc):
 * the whole stage on the given query and reference chunks (host views, as
 * for gta_set_queries / gta_set_references), with every device buffer of the
 * call -- the uploaded chunks, the pair state, backtraces, work lists --
 * carved from one scratch array per device of GTA_NOALLOC_SCRATCH_BYTES
 * that the library holds until the process ends (its address is taken at
 * gta_create).  Results go to the caller's host buffers: alndata nq*nr*15
 * floats, tfms nq*nr*12 floats, alns at least gta_alignment_section_size()
 * bytes; they equal gta_process's.  The plain form runs (no graph, no
 * profiling, no autotuning).  Chunks that need more scratch are rejected
 * with GTA_ERR_SIZE before any result is written.  Calls that overlap on one
 * device race on its scratch (the array keeps the last call's contents; the
 * stage clears what it needs).  The call replaces the context's resident
 * chunks: set them again before the next gta_process. */
#define GTA_NOALLOC_SCRATCH_BYTES (64u << 20)
int gta_process_noalloc(gta_ctx* ctx, const gta_chunk_view* queries, const gta_chunk_view* refs,
                        float* alndata, float* tfms, char* alns, size_t alns_bytes);

/* Same computation without the device-to-host copy; results stay in HBM
 * (for callers that consume them on the device, and for timing the stage
 * with inputs and outputs resident). */
int gta_process_resident(gta_ctx* ctx, double* seconds, double* dp_cells);

/* ---- measurement (no counterpart in the reference, which only times the
 * whole ProcessBlock with a wall clock, CuBatch.cu:450,684) ---------------
 * With profiling on, every kernel launch of gta_process* is bracketed by HIP
 * events on the stage stream; per kernel class the library reports the
 * number of launches, the summed device time and the algorithmic bytes
 * moved (per aligned pair and pass: 24 B covariance, 28 B TM-score,
 * 28 B thresholded covariance; 112 B per slot iteration; DP: 2 bits per
 * cell + 12 B per residue) and the algorithmic FP32 operations (per aligned
 * pair and pass: 24 covariance, 25 thresholded covariance, 33 TM-score;
 * 200 per slot iteration; DP: 12 per cell). */
typedef struct gta_kernel_stat {
    const char* name;
    int launches;
    double seconds;
    double alg_bytes;
    double alg_flops;
} gta_kernel_stat;

int gta_set_profiling(gta_ctx* ctx, int on);
/* fills up to n entries; returns the number of kernel classes */
int gta_kernel_stats(const gta_ctx* ctx, gta_kernel_stat* out, int n);
/* DP cells of the last gta_process* call: useful = sum of Lq x Lr over the
 * fills that ran; launched = the lane-steps the strip sweep executed for them
 * (64-row strips x (Lr + 63) skewed steps rounded up to 16-step blocks). */
int gta_dp_cell_counts(const gta_ctx* ctx, double* useful, double* launched);

/* ---- DP launch-form autotuning (no counterpart in the reference, whose DP
 * tile shape and launch geometry are fixed, dpw_btck.cu:78-615) ------------
 * Each pair's DP fill runs as one wave, as one workgroup of 4, 8 or 16 waves,
 * or pipelined over the whole device.  Off (the default), large pairs take
 * the grid form when they number at most half the compute units and the
 * W-wave form otherwise.  On, pairs of >= 4 strips of 64 rows and >= 2^18
 * cells are grouped by length bucket (log2 of the strip count, log2 of the
 * reference length, log2 of the bucket's pair count); at the first
 * gta_process on a chunk pair, each bucket without a cached choice is timed
 * over the candidate forms (whole-chunk fills, identity superpositions) and
 * takes the fastest.  Choices are kept in the context and, when cache_path is
 * non-NULL, read from that text file at this call and appended to it after
 * tuning (one line per bucket: device tag, bucket, form, milliseconds per
 * candidate).  Results do not depend on the form (the fills are bit-identical). */
int gta_set_autotune(gta_ctx* ctx, int on, const char* cache_path);

/* message of the last error on ctx (thread-local when ctx is NULL) */
const char* gta_last_error(const gta_ctx* ctx);

/* size in bytes of the alignment section for the given chunk shapes
 * (reference: CuMemoryBase::GetSizeOfDP2Alns) */
size_t gta_alignment_section_size(int nqrys, int nqryposs, int nrefs, int nrefposs);

#ifdef __cplusplus
}
#endif

#endif /* GTALIGN_AMD_H */
