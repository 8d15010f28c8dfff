/*
 * hcx_dslim.hpp -- the reference's own C++ entry points of the one-rank
 * search core, implemented over the C ABI of hcx.h (libhcx_dslim.so).
 *
 * Same names, argument meaning and status codes as the reference:
 *   DSLIM_Construct(Index*)                       source/core/include/dslim.h:49
 *   DSLIM_InitializeScorecard(Index*, uint_t)     dslim.h:53
 *   DSLIM_QuerySpectrum(Queries*, Index*, uint_t) dslim.h:184
 *   DSLIM_DeallocateIonIndex(Index*)              dslim.h:149
 *   DSLIM_DeallocateSC()                          dslim.h:164
 *   expeRT::ModelTailFit(Results*)                source/core/include/expeRT.h:143
 *   DFile_PrintScore(...) -- the result sink      source/core/include/dslim_fileout.h:29
 *   DFile_InitFiles(), DFile_DeinitFiles(),        dslim_fileout.h:28-32 (the PSM TSV writer,
 *   DFile_PrintPartials(...)                       dslim_fileout.cpp:45-152)
 * and the parameter block the core reads, `extern gParams params`, defined by
 * the application (the reference defines it in source/app/hicops.cpp:31).
 *
 * The types carry the fields of source/core/include/slm_dsts.h that this path
 * reads or writes, in the same order.  Device state (ion index, scorecard,
 * query buffers) is kept in a side table keyed by Index*: nothing here puts a
 * device pointer into a field the caller frees.
 *
 * DFile_PrintScore is called, as in the reference (dslim_query.cpp:771-808),
 * for every spectrum with cpsms >= params.min_cpsm whose e-value is below
 * params.expect_max.  The library's own definition discards the PSM; an
 * application defines its own DFile_PrintScore to receive them (an
 * executable's definition takes precedence over the shared library's).
 */
#ifndef HCX_DSLIM_HPP
#define HCX_DSLIM_HPP

#include <cstdint>
#include <string>

#include "hcx.h"

typedef int32_t  status_t;
typedef uint32_t uint_t;
typedef int32_t  int_t;
typedef uint16_t ushort_t;
typedef uint64_t ull_t;
typedef char     AA;

/* status codes: source/core/include/slmerr.h:24-43 */
#define SLM_SUCCESS        HCX_SUCCESS
#define ERR_INVLD_PTR      HCX_ERR_INVLD_PTR
#define ERR_INVLD_MEMORY   HCX_ERR_INVLD_MEMORY
#define ERR_INVLD_PARAM    HCX_ERR_INVLD_PARAM
#define ERR_BAD_MEM_ALLOC  HCX_ERR_BAD_MEM_ALLOC
#define ERR_INVLD_SIZE     HCX_ERR_INVLD_SIZE
#define ERR_FILE_NOT_FOUND HCX_ERR_FILE_NOT_FOUND
#define ERR_NOT_ENOUGH_DATA HCX_ERR_NOT_ENOUGH_DATA
#define ERR_INVLD_INDEX    HCX_ERR_INVLD_INDEX
#define ERR_SLM_ALREADY_INIT HCX_ERR_SLM_ALREADY_INIT
#define ERR_INVLD_SORT     HCX_ERR_INVLD_SORT
#define ERR_INVLD_MOD      HCX_ERR_INVLD_MOD
#define ERR_INVLD_NODE_RANK HCX_ERR_INVLD_NODE_RANK
#define ERR_FILE_ERROR     HCX_ERR_FILE_ERROR

/* compile-time knobs of the reference (CMakeLists.txt:157-177) */
#ifndef MAX_HYPERSCORE
#define MAX_HYPERSCORE 100
#endif
#ifndef MAX_SHDPEAKS
#define MAX_SHDPEAKS 80
#endif
#ifndef MAX_MOD_TYPES
#define MAX_MOD_TYPES HCX_MAX_MOD_TYPES
#endif

/* slm_dsts.h:41-71 */
typedef struct _SLM_varAA {
    AA       residues[4];
    uint_t   modMass;          /* (uint)(Da * scale) */
    ushort_t aa_per_peptide;
} SLM_varAA;

/* slm_dsts.h:73-103 */
typedef struct _SLM_Mods {
    ushort_t  vmods_per_pep;
    ushort_t  num_vars;
    SLM_varAA vmods[MAX_MOD_TYPES];
} SLM_vMods;

/* slm_dsts.h:105-120 */
typedef struct _pepSeq {
    AA      *seqs;     /* all sequences of one length, back to back */
    ushort_t peplen;
    uint_t   AAs;      /* number of characters in seqs */
} PepSeqs;

/* slm_dsts.h:121-144 */
typedef struct _modAA {
    ull_t  sites;      /* bit i: residue i carries a variable mod */
    uint_t modNum;     /* 4 bits per mod: mod type + 1, first mod in bits 0-3 */
} modAA;

/* slm_dsts.h:146-282: sorted by Mass within one length */
typedef struct _pepEntry {
    float  Mass;
    uint_t seqID;
    modAA  sites;
} pepEntry;

/* slm_dsts.h:306-320 */
typedef struct _SLMchunk {
    uint_t *iA;
    uint_t *bA;
} SLMchunk;

/* slm_dsts.h:322-352: one per peptide length */
typedef struct _Index {
    uint_t    pepCount;
    uint_t    modCount;
    uint_t    totalCount;
    uint_t    lclpepCnt;
    uint_t    lclmodCnt;
    uint_t    lcltotCnt;       /* entries in pepEntries */
    uint_t    nChunks;
    uint_t    chunksize;
    uint_t    lastchunksize;
    PepSeqs   pepIndex;
    pepEntry *pepEntries;
    SLMchunk *ionIndex;        /* host copy of iA/bA, filled only on request (DSLIM_ExportIonIndex) */
} Index;

/* slm_dsts.h:33-39 */
typedef enum _DistPolicy { _chunk, _cyclic, _zigzag } DistPolicy;

/* slm_dsts.h:354-418: the fields the search core reads */
typedef struct _globalParams {
    uint_t      threads = 1;
    uint_t      maxprepthds = 1;
    uint_t      min_len = 6;
    uint_t      max_len = 40;
    uint_t      maxz = 3;
    uint_t      topmatches = 10;
    uint_t      scale = 100;
    uint_t      min_shp = 4;
    uint_t      min_cpsm = 4;
    uint_t      nodes = 1;
    uint_t      myid = 0;
    uint_t      spadmem = 2048;
    uint_t      min_mass = 500;
    uint_t      max_mass = 5000;
    uint_t      dF = 0;
    int_t       base_int = 100000;
    int_t       min_int = 1000;
    double      dM = 500.0;
    double      res = 0.01;
    double      expect_max = 20.0;
    std::string dbpath;
    std::string datapath;
    std::string workspace;
    std::string modconditions;
    DistPolicy  policy = _cyclic;
    SLM_vMods   vModInfo = {};
} gParams;

/* slm_dsts.h:422-559: CSR spectra of one batch */
typedef struct _queries {
    uint_t *moz;
    uint_t *intensity;
    uint_t *idx;           /* numSpecs + 1 row offsets */
    float  *precurse;
    int_t  *charges;
    float  *rtimes;
    int_t   numPeaks;
    int_t   numSpecs;
    int_t   batchNum;
    int_t   fileNum;
} Queries;

/* slm_dsts.h:562-662: one PSM */
typedef struct _heapEntry {
    ushort_t idxoffset;    /* index of the peptide's length in the Index array */
    ushort_t sharedions;
    ushort_t totalions;
    ushort_t fileIndex;
    int_t    psid;         /* row in that length's pepEntries */
    float    pmass;
    int_t    pchg;
    float    rtime;
    float    hyperscore;
} hCell;

/* the data members of minHeap<hCell> (minheap.h:33-36); the top-K heap of
 * this path lives on the device, so the library never touches these */
typedef struct _hcx_heap_storage {
    int    size;
    int    capacity;
    hCell *array;
} hcx_heap_storage;

/* slm_dsts.h:694-754 (56 bytes, as in the reference) */
typedef struct _Results {
    uint_t  cpsms;
    hcx_heap_storage topK;
    int_t   mu;
    int_t   beta;
    int_t   minhypscore;
    int_t   maxhypscore;
    int_t   nexthypscore;
    double *survival;      /* 2 + 10 * max_hyperscore bins */
} Results;

extern gParams params;

/* The peptide database stage (lbe.h, mods.h): one FASTA file of equal-length
 * peptides per Index.  Device-resident copies of the sequences and entries
 * are kept beside the Index, so DSLIM_Construct after LBE_Initialize builds
 * from them without a second upload (the Index's host arrays are filled as
 * the reference fills them and must not be changed in between). */
status_t MODS_Initialize();                                           /* mods.cpp:64-103          */
status_t LBE_CountPeps(char *filename, Index *index, uint_t explen);  /* lbe_internal.cpp:369-467 */
status_t LBE_CreatePartitions(Index *index);                          /* lbe_internal.cpp:308-352 */
status_t LBE_Initialize(Index *index);                                /* lbe_internal.cpp:119-202 */
status_t LBE_Distribute(Index *index);                                /* lbe_internal.cpp:256-295 */
status_t LBE_Deinitialize(Index *index);                              /* lbe_internal.cpp:241     */

status_t DSLIM_Construct(Index *index);
status_t DSLIM_InitializeScorecard(Index *index, uint_t idxs);
status_t DSLIM_QuerySpectrum(Queries *ss, Index *index, uint_t idxchunk);
status_t DSLIM_DeallocateIonIndex(Index *index);
status_t DSLIM_DeallocateSC();
status_t DFile_PrintScore(Index *index, uint_t specid, float pmass, hCell *psm, double e_x, uint_t npsms);

/* The PSM TSV writer (dslim_fileout.cpp:45-152).  DFile_InitFiles opens
 * params.threads files "<workspace>/<%m.%d.%Y_%H.%M.%S_><myid>_<f>.tsv" with
 * the reference's header; while they are open, the library's DFile_PrintScore
 * appends one line per PSM in the reference's format (std::to_string of each
 * field, an empty mod_info column).  The reference writes a PSM to the file of
 * the OpenMP thread that scored it; here every PSM reaches the calling host
 * thread, so all lines go to file 0 (the other files hold the header only):
 * the files' concatenated lines are the reference's, in spectrum order.
 * File names come from `queryfiles` (the application's list, hicops.cpp:28,
 * indexed by hCell.fileIndex; an empty name if the application defines none
 * or the index is beyond it).                                              */
status_t DFile_InitFiles();
status_t DFile_DeinitFiles();
status_t DFile_PrintPartials(uint_t specid, Results *resPtr);

/* expeRT.h:58-151: one per host thread in the reference (ePtrs) */
class expeRT {
public:
    status_t ModelTailFit(Results *rPtr);
};

/* ---- no reference counterpart ---- */
/* Histogram size and log-factorial table size of the device path: the
 * reference's MAX_HYPERSCORE / MAX_SHDPEAKS compile-time knobs, at run time
 * (call before DSLIM_Construct; defaults are the knobs above). */
status_t DSLIM_SetLimits(uint_t max_hyperscore, uint_t max_shdpeaks);
/* Query kernel form: HCX_KERNEL_BATCHED (default) or HCX_KERNEL_PLAIN. */
status_t DSLIM_SetKernel(int_t kernel);
/* Host copy of the index: allocates index->ionIndex[0].iA / .bA with new[],
 * freed by DSLIM_DeallocateIonIndex. */
status_t DSLIM_ExportIonIndex(Index *index);
/* The device context, for tuning through hcx.h (NULL before the first call). */
hcx_ctx *DSLIM_Context();

#endif /* HCX_DSLIM_HPP */
