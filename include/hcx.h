/*
 * hcx.h -- C ABI of the MI355X-native one-rank HiCOPS search core.
 *
 * The reference exposes its search core as mangled C++ entry points in
 * libhicops-core (source/core/include/dslim.h, expeRT.h).  This header is the
 * device boundary underneath them: plain C, plain pointers and sizes, no
 * C++ or torch types.  Every entry point returns a status code with the same
 * meaning as the reference's status_t (source/core/include/slmerr.h:24-43).
 *
 * Reference interface each entry point replaces:
 *   hcx_create / hcx_destroy      gParams set-up + DSLIM_InitializeScorecard
 *                                 (source/core/dslim.cpp:390-437), DSLIM_DeallocateSC
 *   hcx_index_build               DSLIM_Construct(Index*)            (dslim.h:49, dslim.cpp:46-138)
 *   hcx_index_export              read-back of Index.ionIndex[0].{iA,bA} (slm_dsts.h:306-320)
 *   hcx_index_free                DSLIM_DeallocateIonIndex(Index*)   (dslim.h:149, dslim.cpp:644-691)
 *   hcx_pepdb_create              LBE_CountPeps(char*, Index*, uint_t) (lbe.h:90, lbe_internal.cpp:369-467)
 *                                 with MODS_Initialize / MODS_ModCounter (mods.cpp:64-103, 397-431)
 *   hcx_pepdb_generate            LBE_Initialize(Index*) (lbe.h:45, lbe_internal.cpp:119-202) with
 *                                 LBE_GeneratePeps, MODS_GenerateMods (mods.cpp:448-496)
 *   hcx_pepdb_counts              LBE_CreatePartitions(Index*) (lbe_internal.cpp:308-352)
 *   hcx_index_build_pepdb         DSLIM_Construct(Index*) on the device-resident pepEntries
 *   hcx_queries_upload            hand-over of a Queries batch       (slm_dsts.h:422-559)
 *   hcx_raw_upload, hcx_queries_pick, hcx_queries_pickpeaks
 *                                 MSQuery::pickpeaks (msquery.cpp:509-569) over a batch, as
 *                                 MSQuery::extractbatch calls it per spectrum (msquery.cpp:238-298)
 *   hcx_queries_info, hcx_queries_export  read-back of a Queries batch (slm_dsts.h:422-468)
 *   hcx_search                    DSLIM_QuerySpectrum(Queries*, Index*, uint_t) (dslim.h:184,
 *                                 dslim_query.cpp:515-842) with expeRT::ModelTailFit(Results*)
 *                                 (expeRT.h:143, expeRT.cpp:572-718) and the e-value block
 *                                 (dslim_query.cpp:771-818); results in place of DFile_PrintScore
 *                                 (dslim_fileout.h:29)
 *   hcx_tailfit_batch             expeRT::ModelTailFit(Results*) over a batch of histograms
 *   hcx_part_export               the multi-rank query's partial result (dslim_query.cpp:733-763,
 *                                 expeRT::StoreIResults expeRT.cpp:870-906) of one index shard
 *   hcx_combine                   DSLIM_Score::CombineResults (dslim_score.cpp:198-386) with
 *                                 expeRT::Reconstruct (expeRT.cpp:909-936) and
 *                                 expeRT::ModelTailFit(double&, int) (expeRT.cpp:720-852)
 *   hcx_surrogate_*               no counterpart: learned tail-parameter model with audit fallback
 */
#ifndef HCX_H
#define HCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: same values as source/core/include/slmerr.h:24-43
 * (SLM_ERROR_BASE -3000 minus 1 ... 13) ---- */
#define HCX_SUCCESS               0
#define HCX_ERR_INVLD_MEMORY      (-3001)
#define HCX_ERR_INVLD_PARAM       (-3002)
#define HCX_ERR_INVLD_INDEX       (-3003)
#define HCX_ERR_INVLD_SIZE        (-3004)
#define HCX_ERR_SLM_ALREADY_INIT  (-3005)
#define HCX_ERR_FILE_NOT_FOUND    (-3006)
#define HCX_ERR_BAD_MEM_ALLOC     (-3007)
#define HCX_ERR_INVLD_PTR         (-3008)
#define HCX_ERR_INVLD_SORT        (-3009)
#define HCX_ERR_INVLD_MOD         (-3010)
#define HCX_ERR_INVLD_NODE_RANK   (-3011)
#define HCX_ERR_NOT_ENOUGH_DATA   (-3012)
#define HCX_ERR_FILE_ERROR        (-3013)
#define HCX_ERR_DEVICE         (-3100) /* a HIP call failed; hcx_last_error() has the text */

/* ---- compile-time knobs of the reference, as run-time fields ----
 * gParams (slm_dsts.h:359-418) plus the CMake knobs MAX_HYPERSCORE and
 * MAX_SHDPEAKS (CMakeLists.txt:171-177).  */
#define HCX_MAX_MOD_TYPES 15   /* slm_dsts.h:28 */

typedef struct hcx_params {
    uint32_t maxz;            /* max fragment charge                  (default 3)      */
    uint32_t scale;           /* bins per Da                           (default 100)    */
    uint32_t min_mass;        /* legal precursor mass range (Da)      (500)            */
    uint32_t max_mass;        /*                                        (5000)           */
    uint32_t dF;              /* fragment tolerance in bins             (tol*100, Q1)    */
    uint32_t min_shp;         /* minimum shared peaks                   (4)              */
    uint32_t min_cpsm;        /* minimum candidate PSMs for an e-value  (4)              */
    uint32_t topmatches;      /* heap capacity = 2^floor(log2(t)+0.999)-1 (10 -> 15)     */
    uint32_t max_hyperscore;  /* histogram has 10*max_hyperscore+2 bins (100)            */
    uint32_t max_shdpeaks;    /* lgfact table has max_shdpeaks+1 entries (80)            */
    uint32_t nmods;           /* number of variable-mod types in mod_mass[]                */
    uint32_t mod_mass[HCX_MAX_MOD_TYPES]; /* (uint)(Da*scale), hicops.cpp:661          */
    double   dM;              /* precursor tolerance in Da; < 0 disables the search (Q8)   */
    double   expect_max;      /* report a PSM only if e < expect_max                        */
} hcx_params;

/* pepEntry (slm_dsts.h:146-282): 24 bytes, identical layout. */
typedef struct hcx_pep_entry {
    float    mass;      /* precursor mass (float, UTILS_CalculatePepMass / ModMass)  */
    uint32_t seq_id;    /* row in the per-length sequence table                      */
    uint64_t sites;     /* bit i set: residue i carries a variable mod               */
    uint32_t mod_num;   /* 4 bits per mod: (mod type + 1), first mod in bits 0-3     */
    uint32_t pad_;
} hcx_pep_entry;

/* Queries (slm_dsts.h:422-468): CSR spectra. */
typedef struct hcx_queries_view {
    const uint32_t *moz;        /* peak bins, (uint)(mz*scale)                          */
    const uint32_t *intensity;  /* normalised intensities                              */
    const uint32_t *idx;        /* num_specs+1 row offsets                             */
    const float    *precurse;   /* precursor masses                                    */
    const int32_t  *charges;    /* precursor charges                                   */
    const float    *rtimes;     /* retention times (may be NULL)                       */
    int32_t         num_specs;
    int32_t         file_num;
} hcx_queries_view;

/* Per-spectrum search result: the top hCell (slm_dsts.h:562-662) plus the
 * Results fields (slm_dsts.h:694-754) and the e-value of dslim_query.cpp:786-799. */
typedef struct hcx_result {
    float    hyperscore;    /* top hit hyperscore (float, as hCell.hyperscore)          */
    int32_t  psid;          /* top hit: index into its length's sorted pepEntries      */
    uint16_t idxoffset;     /* top hit: index of its length in the index list          */
    uint16_t sharedions;    /* bc + yc of the top hit                                   */
    uint16_t totalions;     /* speclen of the top hit's length                          */
    uint16_t flags;         /* HCX_RES_* bits                                           */
    uint32_t cpsms;         /* candidate PSMs (shpk >= min_shp and hs > 0)              */
    int32_t  maxhypscore;   /* (int)(hs*10 + 0.5)                                       */
    int32_t  mu;            /* tail-fit slope * 1e6, truncated (Q19)                    */
    int32_t  beta;          /* tail-fit intercept * 1e6, truncated                       */
    int32_t  minhypscore;   /* tail-fit stt1                                             */
    int32_t  nexthypscore;  /* tail-fit end1                                             */
    uint32_t bc;            /* b-ion count of the top hit                               */
    uint32_t yc;            /* y-ion count of the top hit                               */
    double   e_x;           /* e-value = 10^(mu/1e6*maxhyp + beta/1e6) * cpsms          */
} hcx_result;

#define HCX_RES_VALID        0x0001u /* cpsms >= min_cpsm: top hit + e-value valid     */
#define HCX_RES_REPORTED     0x0002u /* e_x < expect_max (DFile_PrintScore would run)  */
#define HCX_RES_TIE_REPLAYED 0x0004u /* max hyperscore tied: heap replayed in order    */
#define HCX_RES_HIST_CLAMP   0x0008u /* a score bin exceeded the histogram (ref. UB)   */
#define HCX_RES_LGF_CLAMP    0x0010u /* bc or yc exceeded max_shdpeaks (ref. UB)       */
#define HCX_RES_FIT_UB       0x0020u /* tail fit hit an undefined path of the reference */
#define HCX_RES_SURROGATE    0x0040u /* e-value from the surrogate model               */
#define HCX_RES_AUDITED      0x0080u /* spectrum was in the surrogate audit sample      */
#define HCX_RES_COUNT_OVF    0x0100u /* a two-pass chunk hit counter neared overflow; the
                                        search re-ran the spectrum on HCX_KERNEL_CURSOR  */
#define HCX_RES_TWO_PASS     0x0200u /* scored by the two-pass kernel (outside the packed
                                        scan's exact range, or HCX_SCAN_TWO_PASS)        */
#define HCX_RES_FIT_HOST     0x0400u /* tail fit recomputed on the host: the device's slope
                                        or intercept (x 1e6) sat within the log10 error
                                        bound of an integer, or a log10 value of a window
                                        threshold, so the host's FP64 path decided it    */

/* Per-spectrum work counters (the algorithmic-bytes inputs). */
typedef struct hcx_stats {
    uint64_t n_ia;      /* iA entries visited inside the precursor windows         */
    uint64_t n_probe;   /* (length, peak, bin) probes that read bA[bin], bA[bin+1]  */
} hcx_stats;

typedef struct hcx_ctx    hcx_ctx;
typedef struct hcx_index  hcx_index;
typedef struct hcx_qbatch hcx_qbatch;

/* Query kernel choice. */
#define HCX_KERNEL_PLAIN   0  /* one workgroup per spectrum, its chunks in order     */
#define HCX_KERNEL_BATCHED 1  /* persistent, work-stealing over (spectrum, length,
                                 chunk) items                                       */
#define HCX_KERNEL_CURSOR  2  /* one workgroup per spectrum on bA/iA alone: binary
                                 searches + cursors (diagnostic, slow)              */

/* E-value model choice. */
#define HCX_EVAL_TAILFIT   0  /* expeRT log-linear tail fit on every spectrum        */
#define HCX_EVAL_SURROGATE 1  /* MLP prediction, tail fit on the audit sample        */

typedef struct hcx_search_opts {
    int32_t  kernel;          /* HCX_KERNEL_*                                       */
    int32_t  evalue;          /* HCX_EVAL_*                                         */
    uint32_t tile;            /* candidates per LDS tile (0 = library default)      */
    uint32_t audit_every;     /* surrogate: tail-fit every n-th spectrum (0 = 64)   */
    double   audit_tol;       /* surrogate: max |rel e error| before fallback       */
    int32_t  use_graph;       /* capture the search sequence in a hipGraph, replayed
                                 while the launch shape is unchanged                 */
    int32_t  reserved;
} hcx_search_opts;

/* ---- context ---- */
int  hcx_create(const hcx_params *params, int device, hcx_ctx **out);
int  hcx_destroy(hcx_ctx *ctx);
const char *hcx_last_error(void);
int  hcx_heap_capacity(uint32_t topmatches);   /* dslim.cpp:422-428 (Q14) */
/* Query tuning (no reference counterpart).  chunk: peptides per LDS chunk
 * (power of two, 256..32768; set before building indices, which record it);
 * slots: scored-peptide slots per chunk pass; blocks_per_cu: persistent grid
 * of the batched kernel.  0 keeps the current value. */
int  hcx_set_tuning(hcx_ctx *ctx, uint32_t chunk, uint32_t slots, uint32_t blocks_per_cu);
/* Further query tuning options (no reference counterpart).
 * HCX_OPT_HITLIST: LDS hit list of the chunked kernels -- the first entries
 *   of each work item are kept in LDS so the exact-count pass need not read
 *   iA again.  < 0: as many as fit (default); 0: none; n > 0: at most n
 *   (rounded down to a multiple of 64).
 * HCX_OPT_GROUP: (peak, bin) probes per run table of the two-pass kernels.
 *   0 (default): as many as the workgroup's LDS share holds beside the other
 *   arrays, up to 2560 and not below 1024; n: 64..2560.  A work item with
 *   more probes walks its entries group by group (12 bytes of LDS per probe).
 *   The packed kernels always take as many as fit (up to 2048).
 * HCX_OPT_SCAN: HCX_SCAN_AUTO (default): the single-pass packed scan for
 *   every spectrum whose counts and intensity sums provably fit its packed
 *   words, the two-pass scan for the rest; HCX_SCAN_TWO_PASS: the two-pass
 *   scan for every spectrum.  Results are identical.
 * HCX_OPT_GRAB: batched kernels: work items per queue grab, at most (1..256,
 *   default 64).
 * HCX_OPT_FORCE_REDO: test hook: besides the spectra whose two-pass counters
 *   neared overflow (HCX_RES_COUNT_OVF), every n-th spectrum is re-scored on
 *   the cursor kernel (0, the default: none).
 * HCX_OPT_FORCE_HOST_FIT: test hook: besides the fits the device lists as
 *   too close to a boundary (HCX_RES_FIT_HOST), every n-th spectrum's tail
 *   fit is recomputed on the host (0, the default: none). */
#define HCX_OPT_HITLIST    1
#define HCX_OPT_GROUP      3
#define HCX_OPT_SCAN       4
#define HCX_OPT_GRAB       5
#define HCX_OPT_FORCE_REDO 6
#define HCX_OPT_FORCE_HOST_FIT 7
#define HCX_SCAN_AUTO      0
#define HCX_SCAN_TWO_PASS  1
int  hcx_set_option(hcx_ctx *ctx, int32_t option, int64_t value);
int  hcx_get_option(const hcx_ctx *ctx, int32_t option, int64_t *value);
int  hcx_get_tuning(const hcx_ctx *ctx, uint32_t *chunk, uint32_t *slots, uint32_t *blocks_per_cu,
                    int32_t *num_cu, int32_t *max_lds);
/* lgfact table the device uses (utils.h:115-131), max_shdpeaks+1 doubles */
int  hcx_lgfact_table(const hcx_ctx *ctx, double *out, uint32_t n);

/* ---- index (one per peptide length; DSLIM_Construct) ---- */
int  hcx_index_build(hcx_ctx *ctx, uint32_t peplen,
                     const char *seqs, uint32_t nseqs,
                     const hcx_pep_entry *entries, uint32_t nentries,
                     hcx_index **out);
int  hcx_index_info(const hcx_index *idx, uint32_t *peplen, uint32_t *speclen,
                    uint32_t *nentries, uint64_t *nions);
int  hcx_index_export(const hcx_index *idx, uint32_t *iA, uint32_t *bA);
/* last build's device time in ms (fragment generation + radix sort + bA scan) */
int  hcx_index_build_ms(const hcx_index *idx, float *ms);
int  hcx_index_free(hcx_index *idx);

/* ---- peptide database: LBE + MODS on the device ----
 * One hcx_pepdb per peptide length, as the reference keeps one Index per
 * length (a FASTA file of equal-length peptides each).  Masses use the
 * context's residue table and mod masses (hcx_params.mod_mass, nmods);
 * min_mass / max_mass are the legality bounds of LBE_CountPeps.          */
#define HCX_POLICY_CHUNK  0   /* DistPolicy _chunk  (slm_dsts.h:33-39)              */
#define HCX_POLICY_CYCLIC 1   /* DistPolicy _cyclic, gParams' default (slm_dsts.h:416) */

typedef struct hcx_lbe_cfg {
    const char *modconditions; /* gParams.modconditions, "limit AAs n [AAs n ...]" as
                                  hicops.cpp:636-666 builds it; NULL or a limit of 0:
                                  no variants.  At most 8 mods per form (4-bit
                                  modNum nibbles), residues A-Z, each named once,
                                  types 1..nmods of the context.                      */
    uint32_t nodes;            /* ranks sharing the database (gParams.nodes)          */
    uint32_t myid;             /* this rank (gParams.myid)                             */
    int32_t  policy;           /* HCX_POLICY_*; nodes > 1 needs HCX_POLICY_CYCLIC     */
    int32_t  sort;             /* gParams.dM > 0: sort by mass (lbe_internal.cpp:184-187) */
} hcx_lbe_cfg;

typedef struct hcx_pepdb hcx_pepdb;
/* LBE_CountPeps (lbe_internal.cpp:369-467) with MODS_Initialize (mods.cpp:64-103)
 * and MODS_ModCounter (mods.cpp:399-432): lines holds nlines sequences of peplen
 * characters back to back (lower case is upper-cased, as std::toupper does); the
 * mass-legal ones are kept in order and their variants counted on the device.
 * Returns when the counts are known. */
int  hcx_pepdb_create(hcx_ctx *ctx, const hcx_lbe_cfg *cfg, uint32_t peplen, const char *lines,
                      uint32_t nlines, hcx_pepdb **out);
/* counts[6]: pepCount, modCount, totalCount, lclpepCnt, lclmodCnt, lcltotCnt
 * (the Index fields, slm_dsts.h:322-357; LBE_CreatePartitions lbe_internal.cpp:308-352) */
int  hcx_pepdb_counts(const hcx_pepdb *db, uint32_t *counts);
/* LBE_Initialize (lbe_internal.cpp:119-202) for n lengths at once: LBE_GeneratePeps
 * (lbe_internal.cpp:204-229), MODS_GenerateMods (mods.cpp:448-496, each sequence's
 * variants in MODS_ModList order, mods.cpp:333-384, then glibc qsort with
 * cmpvarEntries), and std::sort by mass with CmpPepEntries -- the tie order of
 * libstdc++'s introsort, reproduced exactly. */
int  hcx_pepdb_generate(hcx_pepdb *const *dbs, uint32_t n);
/* seqs: pepCount * peplen characters (Index.pepIndex.seqs), entries: lcltotCnt
 * (Index.pepEntries); either may be NULL. */
int  hcx_pepdb_export(const hcx_pepdb *db, char *seqs, hcx_pep_entry *entries);
/* device time of hcx_pepdb_create (the lines' upload over PCIe, then the
 * filter and counts in HBM) and of the hcx_pepdb_generate call that produced
 * db (all its lengths together); any pointer may be NULL */
int  hcx_pepdb_ms(const hcx_pepdb *db, float *upload_ms, float *count_ms, float *generate_ms);
int  hcx_pepdb_free(hcx_pepdb *db);
/* DSLIM_Construct (dslim.cpp:46-138) over a generated database, entries and
 * sequences taken on the device (no host round trip). */
int  hcx_index_build_pepdb(hcx_ctx *ctx, const hcx_pepdb *db, hcx_index **out);
/* The device mass sort alone on n (key, value) pairs: the order libstdc++'s
 * std::sort leaves pepEntries in when compared by key alone (test hook). */
int  hcx_sort_mass(hcx_ctx *ctx, float *keys, uint32_t *vals, uint32_t n);

/* ---- query batches ---- */
int  hcx_queries_upload(hcx_ctx *ctx, const hcx_queries_view *q, hcx_qbatch **out);
int  hcx_queries_free(hcx_qbatch *qb);

/* ---- Queries producer: MSQuery::pickpeaks on the device ----
 * Raw peaks as MSQuery::readspectrum stores them (msquery.cpp:300-412):
 * mz = (uint)(atof(m/z) * scale), intensity = (uint)(atof(I) * 1000).      */
typedef struct hcx_raw_spectra_view {
    const uint32_t *mz;         /* raw peak bins                                        */
    const uint32_t *intensity;  /* raw intensities x1000                                */
    const uint32_t *idx;        /* num_specs+1 row offsets                              */
    const double   *prec_mz;    /* Z-line value as read (spectrum.prec_mz)              */
    const int32_t  *z;          /* Z-line charge (spectrum.Z)                           */
    const double   *rtime;      /* I RTime value (spectrum.rtime); NULL: 0               */
    int32_t         num_specs;
    int32_t         file_num;
} hcx_raw_spectra_view;

/* pickpeaks knobs: qalen = QALEN (config.hpp, at most peaks kept per spectrum,
 * 1..4096), base_int and min_int = gParams.base_int / min_int
 * (slm_dsts.h:406-407, hicops.cpp:606-610).                                  */
typedef struct hcx_pick_params {
    uint32_t qalen;
    uint32_t base_int;
    uint32_t min_int;
    uint32_t pad_;
} hcx_pick_params;

typedef struct hcx_rawbatch hcx_rawbatch;

/* Raw spectra to device memory (the device copy is never modified, so one
 * raw batch can be picked any number of times).                           */
int  hcx_raw_upload(hcx_ctx *ctx, const hcx_raw_spectra_view *raw, hcx_rawbatch **out);
int  hcx_raw_free(hcx_rawbatch *rb);
/* pickpeaks over every spectrum of rb (msquery.cpp:509-569), bit-exact: the
 * reference's quicksort (keyval.h:48-132) and its order of equal intensities,
 * scaling to base_int, the min_int cut, the last speclen peaks in ascending-
 * intensity order; precurse = (float)prec_mz, charges = MAX(1, z), rtimes =
 * (float)MAX(0.0, rtime).  The batch stays in device memory, ready for
 * hcx_search.  kernel_ms / total_ms (optional): device time of the
 * normalisation kernel and of the whole pass (HIP events).                  */
int  hcx_queries_pick(hcx_ctx *ctx, hcx_rawbatch *rb, const hcx_pick_params *pp, hcx_qbatch **out,
                      float *kernel_ms, float *total_ms);
/* hcx_raw_upload + hcx_queries_pick + hcx_raw_free.                        */
int  hcx_queries_pickpeaks(hcx_ctx *ctx, const hcx_raw_spectra_view *raw, const hcx_pick_params *pp,
                           hcx_qbatch **out);
/* Size and host copy of a batch (any argument may be NULL): moz/intensity
 * npeaks entries, idx num_specs+1, precurse/charges/rtimes num_specs.       */
int  hcx_queries_info(const hcx_qbatch *qb, int32_t *num_specs, uint64_t *npeaks);
int  hcx_queries_export(const hcx_qbatch *qb, uint32_t *moz, uint32_t *intensity, uint32_t *idx,
                        float *precurse, int32_t *charges, float *rtimes);
/* The batch facts the search sizes its exact scan by (no reference
 * counterpart; computed on the device at upload or pick time): the largest
 * spectrum, the largest intensity, the most peaks of one spectrum within any
 * 2*dF+1 consecutive bins, and per spectrum that count (mult) and the largest
 * intensity sum within 2*dF+1 bins (wsum, saturating at 2^32-1): a spectrum
 * is scored by the packed single-pass scan when (speclen/2) * mult < 256 and
 * (speclen/2) * wsum < 2^24 for the longest length.  mult, wsum: num_specs
 * entries.  Any argument may be NULL.                                        */
int  hcx_queries_facts(const hcx_qbatch *qb, uint32_t *max_peaks, uint32_t *max_int, uint32_t *max_mult,
                       uint32_t *mult, uint32_t *wsum);

/* ---- search: DSLIM_QuerySpectrum + ModelTailFit + e-value ----
 * results:  num_specs entries (host).            Required.
 * hist:     num_specs * (10*max_hyperscore+2) u32 (host).  Optional (NULL).
 * stats:    num_specs entries (host).            Optional (NULL).          */
int  hcx_search(hcx_ctx *ctx, hcx_index *const *idx, int nidx, hcx_qbatch *qb,
                const hcx_search_opts *opts, hcx_result *results,
                uint32_t *hist, hcx_stats *stats);

/* Autotuner (no reference counterpart): times the scan of `kernel`
 * (HCX_KERNEL_BATCHED or _PLAIN) on this batch for each candidate, keeps the
 * fastest in the context and returns it.  With HCX_SCAN_AUTO the candidates
 * are the batched kernels' grab sizes 8..256 (the packed kernel's LDS layout
 * is fixed by the chunk size); with HCX_SCAN_TWO_PASS, slots
 * 512/1024/1536/2048 x workgroups per CU 1/2/3 (x grab 32/64/128 for the
 * batched form).  The chunk size is the indices' own (they are built for
 * it).  Runs 2 searches per candidate. */
typedef struct hcx_tuning {
    uint32_t chunk;
    uint32_t slots;
    uint32_t blocks_per_cu;
    int32_t  hitlist;
    int32_t  grab;          /* batched: work items per queue grab */
    float    kernel_ms;     /* the chosen candidate's scan time */
} hcx_tuning;
int  hcx_autotune(hcx_ctx *ctx, hcx_index *const *idx, int nidx, hcx_qbatch *qb, int32_t kernel,
                  hcx_tuning *best);

/* Device-resident variant for timing: runs the search and leaves results on
 * the device; hcx_search_fetch copies them out.  Returns the query kernel's
 * device time (HIP events on the library stream) in *kernel_ms (with
 * use_graph: the replayed graph's time). */
int  hcx_search_device(hcx_ctx *ctx, hcx_index *const *idx, int nidx, hcx_qbatch *qb,
                       const hcx_search_opts *opts, float *kernel_ms, float *total_ms);
int  hcx_search_fetch(hcx_ctx *ctx, hcx_qbatch *qb, hcx_result *results,
                      uint32_t *hist, hcx_stats *stats);
/* Device times (HIP events) of the last hcx_search_device: the scan kernel,
 * the e-value stage (the plain tail fit, or the surrogate's features +
 * network + e-values, without its audit), and the whole search.  A search
 * replayed from a hipGraph is timed as a whole: its scan and e-value times
 * are the graph's.                                                          */
int  hcx_last_timing(hcx_ctx *ctx, float *scan_ms, float *evalue_ms, float *total_ms);
/* The last timed search's scan split by kernel (HIP events): the packed
 * kernel, and the two-pass kernel over the spectra the packed one leaves (or
 * over all, with HCX_SCAN_TWO_PASS).  Either is 0 when it did not run.      */
int  hcx_last_scan_split(hcx_ctx *ctx, float *packed_ms, float *two_pass_ms);

/* Allocation-free search: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same search as hcx_search
 * (HCX_EVAL_TAILFIT, the context's tuning) with no hipMalloc, hipFree or
 * other allocating call.  The caller passes the inputs (indices, batch) and
 * device outputs: d_results (num_specs), d_hist (num_specs * (10 *
 * max_hyperscore + 2) u32, required: the tail fit reads it) and d_stats
 * (num_specs, or NULL).  The scratch is a static device array of 64 MiB that
 * every GPU holds until the process ends; a search whose scratch exceeds it
 * (about 70 + 8 x indices + (index chunks) / 16 bytes per spectrum, the
 * rest holding the work items' best scores) returns HCX_ERR_INVLD_SIZE.  All calls on one GPU share that GPU's array: calls
 * that overlap on one GPU (from several threads or contexts) race.  Returns
 * when the results are in d_results.                                        */
int  hcx_search_noalloc(hcx_ctx *ctx, hcx_index *const *idx, int nidx, hcx_qbatch *qb, int32_t kernel,
                        hcx_result *d_results, uint32_t *d_hist, hcx_stats *d_stats);

/* Candidate dump for parity tests: every candidate with shpk >= min_shp
 * and hs > 0, as (spectrum, idxoffset, psid, bc, yc, ibc, iyc, hs).
 * Returns the count in *n (capacity cap).                                   */
typedef struct hcx_cand {
    uint32_t spec;
    uint16_t idxoffset;
    uint16_t bc;
    int32_t  psid;
    uint16_t yc;
    uint16_t pad_;
    uint64_t ibc;
    uint64_t iyc;
    float    hyperscore;
    uint32_t pad2_;
} hcx_cand;
int  hcx_search_candidates(hcx_ctx *ctx, hcx_index *const *idx, int nidx, hcx_qbatch *qb,
                           hcx_cand *out, uint64_t cap, uint64_t *n);

/* ---- expeRT::ModelTailFit over a batch of histograms ----
 * hist: n * hbins u32 counts; cpsms, maxhyp: n; out: mu, beta, minhyp, nexthyp. */
int  hcx_tailfit_batch(hcx_ctx *ctx, const uint32_t *hist, uint32_t hbins, uint32_t n,
                       const uint32_t *cpsms, const int32_t *maxhyp,
                       int32_t *mu, int32_t *beta, int32_t *minhyp, int32_t *nexthyp,
                       uint32_t *flags);

/* ---- partial results of index shards and their merge ----
 * A rank of the reference searches its share of the database and writes, per
 * spectrum, a partRes and the u16 counts of its score histogram from the
 * first to the last occupied bin (slm_dsts.h:757-854, 913-950); one rank
 * then merges the shares (DSLIM_Score::CombineResults).  Here the shards are
 * index sets searched one after another on one GPU.                         */
typedef struct hcx_part_res {   /* partRes, slm_dsts.h:757-854 */
    uint16_t min;    /* first occupied histogram bin */
    uint16_t max2;   /* last occupied histogram bin */
    float    max;    /* the top hit's hyperscore */
    int32_t  N;      /* candidates scored (cpsms) */
    int32_t  qID;    /* spectrum id */
} hcx_part_res;
typedef struct hcx_fresult {    /* fResult, slm_dsts.h:869-905 */
    int32_t eValue;  /* e-value x 1e6, truncated */
    int32_t specID;  /* the winning shard's spectrum id; -1: not reported */
    int32_t npsms;   /* candidates over all shards; 0: not reported */
} hcx_fresult;
/* The partial result of the indices last searched with qb (any kernel, any
 * e-value method): spectrum s gets qID = qid0 + s; spectra with no candidate
 * get a zero hcx_part_res.  The counts of bins min..max2 go to slices + s *
 * slot (the rest of the slot zero); above 65500 candidates each count k is
 * stored as (u16)((double)(int)(k * 65500) / cpsms) with the int product
 * wrapping as the reference's does (expeRT.cpp:880-890).  A histogram with
 * more than `slot` occupied-range bins is HCX_ERR_INVLD_SIZE (the reference,
 * with slot = psize = 128, writes past its spectrum's slot).  packs: nspec,
 * slices: nspec * slot u16, host memory. */
int  hcx_part_export(hcx_ctx *ctx, hcx_qbatch *qb, int32_t qid0, uint32_t slot, hcx_part_res *packs,
                     uint16_t *slices);
/* The merge of nshards partial results of the same nspec spectra, shards in
 * order: counts summed into one histogram per spectrum (decoded from 65500
 * levels above 65500 candidates), the shard of the highest hyperscore as the
 * key (strictly greater: the first shard wins a tie), and for spectra with
 * a key and >= min_cpsm candidates the tail fit of the sum with its slope and
 * intercept kept in FP64; e < expect_max is reported (eValue = e x 1e6,
 * specID = the key shard's qID, npsms, keys[s] = key), anything else gets
 * eValue = expect_max x 1e6, specID -1, npsms 0, keys[s] = nshards.  A fit
 * whose window or truncation lies within the device log10's error bound of
 * a boundary is done again on the host.  Host memory. */
int  hcx_combine(hcx_ctx *ctx, uint32_t nshards, uint32_t nspec, uint32_t slot, const hcx_part_res *const *packs,
                 const uint16_t *const *slices, hcx_fresult *out, int32_t *keys);

/* ---- learned tail-parameter surrogate (no reference counterpart) ----
 * A 128-128-128-3 MLP (BF16 MFMA, FP32 accumulate) maps a spectrum's score
 * histogram (its log-survival function at 124 points below the top score)
 * to the window of expeRT::ModelTailFit's line fit: the first bin at or
 * below 22 % of the log-survival drop and the last at or above 87 %
 * (expeRT.cpp:657-681).  It is trained on the windows the plain fit
 * (expeRT::ModelTailFit) chose.  With opts.evalue = HCX_EVAL_SURROGATE,
 * hcx_search tests the 64 bins around each predicted end for the exact
 * threshold crossing; where both are found the line is fitted through that
 * window (HCX_RES_SURROGATE), which gives the plain fit's ints; elsewhere
 * the spectrum takes the plain fit.  The plain fit also runs on every
 * audit_every-th spectrum (0: 64) and replaces the values there
 * (HCX_RES_AUDITED); if any audited e-value differs by more than audit_tol
 * (relative; 0: 1e-3) the whole batch falls back to the plain fit.
 * hcx_surrogate_last_audit reports what happened.                         */
typedef struct hcx_audit {
    uint32_t audited;       /* spectra with a plain fit in the sample          */
    uint32_t disagreed;     /* of those, relative e-value error > audit_tol     */
    double   max_rel_err;   /* largest relative error in the sample            */
    int32_t  fell_back;     /* 1: the batch got plain e-values throughout      */
    uint32_t located;       /* spectra whose predicted window was located      */
    uint32_t plain_fit;     /* spectra the prediction missed: plain fit        */
    uint32_t pad_;
} hcx_audit;
/* number of float parameters (weights + biases) of the model */
int  hcx_surrogate_param_count(void);
/* He-uniform initial weights from a seed; targets unstandardised */
int  hcx_surrogate_init(hcx_ctx *ctx, uint64_t seed);
/* parameters in and out: n = hcx_surrogate_param_count() floats, plus the
 * target standardisation (mean[3], std[3]) */
int  hcx_surrogate_set(hcx_ctx *ctx, const float *params, uint32_t n, const float *norm6);
int  hcx_surrogate_get(hcx_ctx *ctx, float *params, uint32_t n, float *norm6);
/* train on a batch last searched with HCX_EVAL_TAILFIT (its histograms and
 * fits are on the device): `steps` Adam steps over mini-batches of `batch`
 * spectra.  *loss: mean standardised squared error over the last 10 % of steps. */
int  hcx_surrogate_train(hcx_ctx *ctx, hcx_qbatch *qb, uint32_t steps, uint32_t batch, float lr,
                         uint64_t seed, float *loss);
/* one forward + backward pass with the current parameters: X n x 128
 * features (rounded to bf16 on the device), T n x 4 (three standardised
 * targets + weight), n a multiple of 16; grad receives dL/dparams.       */
int  hcx_surrogate_gradient(hcx_ctx *ctx, const float *X, const float *T, uint32_t n,
                            float *grad, float *loss);
int  hcx_surrogate_last_audit(const hcx_ctx *ctx, hcx_audit *out);
/* lane-layout self-check of the BF16 MFMA tile: D = A B, 16 x 16 row-major */
int  hcx_selftest_mfma(hcx_ctx *ctx, const float *A, const float *B, float *D);

/* log2 of the integers 1 .. n as the packed scan's score estimate computes
 * it (v_log_f32), and the error bound it assumes (*bound): a test checks the
 * first against log2 over the whole range the estimate sees (n = 2^24 + 1). */
int  hcx_selftest_log2(hcx_ctx *ctx, uint32_t n, float *out, double *bound);

/* ---- host self-test of the raw / speclen divider the query kernels use ---- */
int  hcx_selftest_divider(uint32_t d, const uint32_t *nums, uint32_t count, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif /* HCX_H */
