// dslim_hcx.cpp -- the reference's DSLIM construct / query / e-value entry
// points (include/hcx_dslim.hpp) over the C ABI of libhcx.
//
// One device context per process, created from `params` on first use, as
// the reference builds its scorecard from the global parameter block
// (dslim.cpp:390-437).  Device indices live in a side table keyed by Index*.
#include "../../include/hcx_dslim.hpp"

#include <algorithm>
#include <cstddef>
#include <ctime>
#include <fstream>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

// layouts an application built against source/core/include/slm_dsts.h relies on
static_assert(sizeof(pepEntry) == 24, "pepEntry layout");
static_assert(sizeof(hCell) == 28, "hCell layout");
static_assert(sizeof(Results) == 56 && offsetof(Results, mu) == 24 && offsetof(Results, survival) == 48,
              "Results layout");
static_assert(sizeof(pepEntry) == sizeof(hcx_pep_entry), "pepEntry and hcx_pep_entry");

namespace {

struct PepDbRef {
    hcx_pepdb *db = nullptr;
    const pepEntry *entries = nullptr;   // the host array LBE_Initialize filled
};

struct State {
    std::mutex mu;
    hcx_ctx *ctx = nullptr;
    std::unordered_map<const Index *, hcx_index *> idx;
    std::unordered_map<const Index *, PepDbRef> pdb;   // LBE_CountPeps .. LBE_Deinitialize
    uint_t max_hyperscore = MAX_HYPERSCORE;
    uint_t max_shdpeaks = MAX_SHDPEAKS;
    int_t kernel = HCX_KERNEL_BATCHED;
    uint_t spectrumID = 0;   // running spectrum number across batches (dslim_query.cpp:808-812)
};

State &st() {
    static State s;
    return s;
}

// gParams -> hcx_params (the fields the search core reads)
status_t make_params(hcx_params *p) {
    *p = hcx_params{};
    p->maxz = params.maxz;
    p->scale = params.scale;
    p->min_mass = params.min_mass;
    p->max_mass = params.max_mass;
    p->dF = params.dF;
    p->min_shp = params.min_shp;
    p->min_cpsm = params.min_cpsm;
    p->topmatches = params.topmatches;
    p->max_hyperscore = st().max_hyperscore;
    p->max_shdpeaks = st().max_shdpeaks;
    if (params.vModInfo.num_vars > HCX_MAX_MOD_TYPES) return ERR_INVLD_PARAM;
    p->nmods = params.vModInfo.num_vars;
    for (uint_t i = 0; i < p->nmods; i++) p->mod_mass[i] = params.vModInfo.vmods[i].modMass;
    p->dM = params.dM;
    p->expect_max = params.expect_max;
    return SLM_SUCCESS;
}

// caller holds st().mu
status_t ensure_ctx() {
    State &s = st();
    if (s.ctx) return SLM_SUCCESS;
    hcx_params p;
    status_t rc = make_params(&p);
    if (rc != SLM_SUCCESS) return rc;
    return hcx_create(&p, 0, &s.ctx);
}

}  // namespace

// The application's list of query files (hicops.cpp:28).  Weak, so that
// applications without one still link.
extern std::vector<std::string> queryfiles __attribute__((weak));

namespace {

// dslim_fileout.cpp:29-33: the open TSV files
struct TsvFiles {
    std::mutex mu;
    std::ofstream *tsvs = nullptr;
    uint_t n = 0;
};
TsvFiles &tsv() {
    static TsvFiles t;
    return t;
}

// dslim_fileout.cpp:160-176
std::string datetime() {
    time_t raw;
    std::time(&raw);
    struct tm tmv;
    localtime_r(&raw, &tmv);
    char buf[80];
    std::strftime(buf, sizeof buf, "%m.%d.%Y_%H.%M.%S_", &tmv);
    return std::string(buf) + std::to_string(params.myid);
}

}  // namespace

// dslim_fileout.cpp:45-76
status_t DFile_InitFiles() {
    TsvFiles &t = tsv();
    std::lock_guard<std::mutex> g(t.mu);
    if (t.tsvs) return SLM_SUCCESS;
    const uint_t n = params.threads ? params.threads : 1;
    t.tsvs = new (std::nothrow) std::ofstream[n];
    if (!t.tsvs) return ERR_BAD_MEM_ALLOC;
    t.n = n;
    const std::string common = params.workspace + '/' + datetime();
    for (uint_t f = 0; f < n; f++) {
        t.tsvs[f].open(common + "_" + std::to_string(f) + ".tsv");
        t.tsvs[f] << "file\t" << "scan_num\t" << "prec_mass\t" << "charge\t"
                  << "retention_time\t" << "peptide\t" << "matched_ions\t"
                  << "total_ions\t" << "calc_pep_mass\t" << "mass_diff\t"
                  << "mod_info\t" << "hyperscore\t" << "expectscore\t"
                  << "num_hits" << std::endl;
    }
    return SLM_SUCCESS;
}

// dslim_fileout.cpp:78-88
status_t DFile_DeinitFiles() {
    TsvFiles &t = tsv();
    std::lock_guard<std::mutex> g(t.mu);
    for (uint_t i = 0; i < t.n; i++) t.tsvs[i].close();
    delete[] t.tsvs;
    t.tsvs = nullptr;
    t.n = 0;
    return SLM_SUCCESS;
}

// dslim_fileout.cpp:90-105
status_t DFile_PrintPartials(uint_t specid, Results *resPtr) {
    if (!resPtr) return ERR_INVLD_PTR;
    TsvFiles &t = tsv();
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.tsvs) return ERR_INVLD_PARAM;
    std::ofstream &o = t.tsvs[0];
    o << std::to_string(specid + 1);
    o << '\t' << std::to_string(resPtr->cpsms);
    o << '\t' << std::to_string(resPtr->mu);
    o << '\t' << std::to_string(resPtr->beta);
    o << '\t' << std::to_string(resPtr->minhypscore);
    o << '\t' << std::to_string(resPtr->nexthypscore);
    o << std::endl;
    return SLM_SUCCESS;
}

// dslim_fileout.cpp:107-152.  While no files are open the PSM is discarded;
// an application that defines DFile_PrintScore receives every reported PSM
// instead.
__attribute__((weak)) status_t DFile_PrintScore(Index *index, uint_t specid, float pmass, hCell *psm, double e_x,
                                                uint_t npsms) {
    TsvFiles &t = tsv();
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.tsvs) return SLM_SUCCESS;
    if (!index || !psm) return ERR_INVLD_PTR;
    const Index *lcl = index + psm->idxoffset;
    const int_t peplen = lcl->pepIndex.peplen;
    const int_t pepid = psm->psid;
    const std::string pep(lcl->pepIndex.seqs + (size_t)lcl->pepEntries[pepid].seqID * peplen, (size_t)peplen);
    const std::vector<std::string> *qf = &queryfiles;
    const std::string file = qf && psm->fileIndex < qf->size() ? (*qf)[psm->fileIndex] : std::string();
    const float mass = lcl->pepEntries[pepid].Mass;
    std::ofstream &o = t.tsvs[0];
    o << file;
    o << '\t' << std::to_string(specid + 1);
    o << '\t' << std::to_string(pmass);
    o << '\t' << std::to_string(psm->pchg);
    o << '\t' << std::to_string(psm->rtime);
    o << '\t' << pep;
    o << '\t' << std::to_string(psm->sharedions);
    o << '\t' << std::to_string(psm->totalions);
    o << '\t' << std::to_string(mass);
    o << '\t' << std::to_string((pmass - mass));
    o << '\t';
    o << '\t' << std::to_string(psm->hyperscore);
    o << '\t' << std::to_string(e_x);
    o << '\t' << std::to_string(npsms);
    o << std::endl;
    return SLM_SUCCESS;
}

status_t DSLIM_SetLimits(uint_t max_hyperscore, uint_t max_shdpeaks) {
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    if (s.ctx) return ERR_INVLD_PARAM;   // fixed once the context exists
    if (max_hyperscore == 0 || max_shdpeaks == 0) return ERR_INVLD_PARAM;
    s.max_hyperscore = max_hyperscore;
    s.max_shdpeaks = max_shdpeaks;
    return SLM_SUCCESS;
}

status_t DSLIM_SetKernel(int_t kernel) {
    if (kernel != HCX_KERNEL_BATCHED && kernel != HCX_KERNEL_PLAIN && kernel != HCX_KERNEL_CURSOR)
        return ERR_INVLD_PARAM;
    st().kernel = kernel;
    return SLM_SUCCESS;
}

hcx_ctx *DSLIM_Context() { return st().ctx; }

namespace {

// DistPolicy -> the C ABI's policy codes
int32_t policy_code() { return params.policy == _chunk ? HCX_POLICY_CHUNK : HCX_POLICY_CYCLIC; }

}  // namespace

// mods.cpp:64-103: the mod conditions are parsed (and checked) where they are
// used, by hcx_pepdb_create; here only that params.modconditions reads.
status_t MODS_Initialize() {
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    status_t rc = ensure_ctx();
    if (rc != SLM_SUCCESS) return rc;
    const std::string mc = params.modconditions.empty() ? std::string("0 M 0") : params.modconditions;
    hcx_lbe_cfg cfg = {mc.c_str(), 1, 0, HCX_POLICY_CYCLIC, 1};
    const char probe[] = "GGGGGGGGGK";
    hcx_pepdb *db = nullptr;
    rc = hcx_pepdb_create(s.ctx, &cfg, 10, probe, 1, &db);
    if (db) hcx_pepdb_free(db);
    return rc;
}

// lbe_internal.cpp:369-467: read one length's peptides (lines not starting
// with '>', a trailing '\r' dropped), keep the mass-legal ones and count their
// variants (on the device).  A line of another length is ERR_INVLD_SIZE, as
// in the reference (which reads an empty line out of range instead).
status_t LBE_CountPeps(char *filename, Index *index, uint_t explen) {
    if (!filename || !index) return ERR_INVLD_PTR;
    if (explen < 2 || explen > 60) return ERR_INVLD_SIZE;
    std::ifstream f(filename);
    if (!f.is_open()) return ERR_INVLD_PARAM;
    std::vector<char> lines;
    std::string line;
    uint_t n = 0;
    while (std::getline(f, line)) {
        if (!line.empty() && line[0] == '>') continue;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() != explen) return ERR_INVLD_SIZE;
        lines.insert(lines.end(), line.begin(), line.end());
        n++;
    }
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    status_t rc = ensure_ctx();
    if (rc != SLM_SUCCESS) return rc;
    const std::string mc = params.modconditions.empty() ? std::string("0 M 0") : params.modconditions;
    hcx_lbe_cfg cfg = {mc.c_str(), params.nodes ? params.nodes : 1, params.myid, policy_code(), params.dM > 0.0};
    hcx_pepdb *db = nullptr;
    rc = hcx_pepdb_create(s.ctx, &cfg, explen, lines.data(), n, &db);
    if (rc != SLM_SUCCESS) return rc;
    auto it = s.pdb.find(index);
    if (it != s.pdb.end()) hcx_pepdb_free(it->second.db);
    s.pdb[index] = PepDbRef{db, nullptr};
    uint32_t c[6];
    hcx_pepdb_counts(db, c);
    index->pepCount = c[0];
    index->modCount = c[1];
    index->totalCount = c[2];
    index->pepIndex.peplen = (ushort_t)explen;
    index->pepIndex.AAs = c[0] * explen;
    return SLM_SUCCESS;
}

// lbe_internal.cpp:308-352
status_t LBE_CreatePartitions(Index *index) {
    if (!index) return ERR_INVLD_PTR;
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.pdb.find(index);
    if (it == s.pdb.end()) return ERR_INVLD_PARAM;   // LBE_CountPeps first
    uint32_t c[6];
    hcx_pepdb_counts(it->second.db, c);
    index->lclpepCnt = c[3];
    index->lclmodCnt = c[4];
    index->lcltotCnt = c[5];
    return SLM_SUCCESS;
}

// lbe_internal.cpp:119-202: the entries of this rank's share, sorted by mass
// when params.dM > 0, into new host arrays as LBE_AllocateMem makes them.
status_t LBE_Initialize(Index *index) {
    if (!index) return ERR_INVLD_PTR;
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.pdb.find(index);
    if (it == s.pdb.end()) return ERR_INVLD_PARAM;
    hcx_pepdb *db = it->second.db;
    uint32_t c[6];
    hcx_pepdb_counts(db, c);
    if (c[0] == 0) return ERR_INVLD_PARAM;             // lbe_internal.cpp:134-141
    status_t rc = hcx_pepdb_generate(&db, 1);
    if (rc != SLM_SUCCESS) return rc;
    AA *seqs = new (std::nothrow) AA[(size_t)c[0] * index->pepIndex.peplen];
    pepEntry *ents = new (std::nothrow) pepEntry[c[5] ? c[5] : 1];
    if (!seqs || !ents) {
        delete[] seqs;
        delete[] ents;
        return ERR_BAD_MEM_ALLOC;
    }
    rc = hcx_pepdb_export(db, seqs, reinterpret_cast<hcx_pep_entry *>(ents));
    if (rc != SLM_SUCCESS) {
        delete[] seqs;
        delete[] ents;
        return rc;
    }
    index->pepIndex.seqs = seqs;
    index->pepEntries = ents;
    index->lclpepCnt = c[3];
    index->lclmodCnt = c[4];
    index->lcltotCnt = c[5];
    it->second.entries = ents;
    return SLM_SUCCESS;
}

// lbe_internal.cpp:256-295: the reference's chunk arithmetic (MAX_IONS
// 0x80000000 ions per chunk, a scorecard of spadmem bytes per thread of
// 24-byte BYC cells).  The device index keeps one chunk per length whatever
// these fields say.
status_t LBE_Distribute(Index *index) {
    if (!index) return ERR_INVLD_PTR;
    const uint_t N = index->lcltotCnt;
    const uint_t speclen = (index->pepIndex.peplen - 1) * params.maxz * 2;
    if (speclen == 0 || N == 0) return ERR_INVLD_PARAM;
    const uint_t maxchunk = 0x80000000u / speclen;
    const uint_t maxchunk2 = params.spadmem / (24u * (params.threads ? params.threads : 1));
    uint_t chunksize = std::min(N, maxchunk);
    chunksize = std::min(chunksize, maxchunk2);
    if (chunksize == 0) return ERR_INVLD_PARAM;
    uint_t nchunks = N / chunksize + ((N % chunksize) != 0);
    index->nChunks = nchunks;
    index->chunksize = chunksize;
    index->lastchunksize = (N % chunksize == 0) ? chunksize : N - chunksize * (N / chunksize);
    return SLM_SUCCESS;
}

// lbe_internal.cpp:241 (DSLIM_Deinitialize, dslim.cpp:644-720): the ion
// index, the peptide arrays and the device database of this Index.
status_t LBE_Deinitialize(Index *index) {
    if (!index) return ERR_INVLD_PTR;
    status_t rc = DSLIM_DeallocateIonIndex(index);
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.pdb.find(index);
    if (it != s.pdb.end()) {
        hcx_pepdb_free(it->second.db);
        s.pdb.erase(it);
    }
    delete[] index->pepEntries;
    index->pepEntries = nullptr;
    delete[] index->pepIndex.seqs;
    index->pepIndex.seqs = nullptr;
    index->pepCount = index->modCount = index->totalCount = 0;
    index->pepIndex.AAs = 0;
    index->pepIndex.peplen = 0;
    return rc;
}

// dslim.cpp:46-138: build the ion index of one peptide length from its
// sequences and mass-sorted pepEntries.
status_t DSLIM_Construct(Index *index) {
    if (!index) return ERR_INVLD_PTR;
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    status_t rc = ensure_ctx();
    if (rc != SLM_SUCCESS) return rc;
    const uint_t L = index->pepIndex.peplen;
    const uint_t n = index->lcltotCnt;
    if (L == 0) return ERR_INVLD_PARAM;
    if (n && (!index->pepEntries || !index->pepIndex.seqs)) return ERR_INVLD_PTR;
    std::vector<hcx_pep_entry> e(n);
    for (uint_t i = 0; i < n; i++) {
        e[i].mass = index->pepEntries[i].Mass;
        e[i].seq_id = index->pepEntries[i].seqID;
        e[i].sites = index->pepEntries[i].sites.sites;
        e[i].mod_num = index->pepEntries[i].sites.modNum;
        e[i].pad_ = 0;
    }
    auto it = s.idx.find(index);
    if (it != s.idx.end()) {
        hcx_index_free(it->second);
        s.idx.erase(it);
    }
    hcx_index *ix = nullptr;
    auto pd = s.pdb.find(index);
    if (pd != s.pdb.end() && pd->second.entries == index->pepEntries && params.dM > 0.0)
        rc = hcx_index_build_pepdb(s.ctx, pd->second.db, &ix);   // the entries LBE_Initialize left on the device
    else
        rc = hcx_index_build(s.ctx, L, index->pepIndex.seqs, index->pepIndex.AAs / L, e.data(), n, &ix);
    if (rc != SLM_SUCCESS) return rc;
    s.idx[index] = ix;
    return SLM_SUCCESS;
}

// dslim.cpp:390-437: the scorecard (LDS counters, histograms, heaps) lives on
// the device and is sized per search; only the context is needed here.
status_t DSLIM_InitializeScorecard(Index *index, uint_t idxs) {
    if (!index && idxs) return ERR_INVLD_PTR;
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    return ensure_ctx();
}

// dslim_query.cpp:515-842: search one batch against the indices
// index[0 .. idxchunk) and report each spectrum's top PSM through
// DFile_PrintScore when cpsms >= min_cpsm and e < expect_max.
status_t DSLIM_QuerySpectrum(Queries *ss, Index *index, uint_t idxchunk) {
    if (!ss || (!index && idxchunk)) return ERR_INVLD_PTR;
    State &s = st();
    std::vector<hcx_index *> ix(idxchunk);
    hcx_ctx *ctx;
    int_t kernel;
    uint_t first;
    {
        std::lock_guard<std::mutex> g(s.mu);
        status_t rc = ensure_ctx();
        if (rc != SLM_SUCCESS) return rc;
        for (uint_t i = 0; i < idxchunk; i++) {
            auto it = s.idx.find(&index[i]);
            if (it == s.idx.end()) return ERR_INVLD_PARAM;   // DSLIM_Construct not called
            ix[i] = it->second;
        }
        ctx = s.ctx;
        kernel = s.kernel;
        first = s.spectrumID;
        s.spectrumID += (uint_t)(ss->numSpecs > 0 ? ss->numSpecs : 0);
    }
    if (ss->numSpecs <= 0) return SLM_SUCCESS;
    hcx_queries_view v;
    v.moz = ss->moz;
    v.intensity = ss->intensity;
    v.idx = ss->idx;
    v.precurse = ss->precurse;
    v.charges = ss->charges;
    v.rtimes = ss->rtimes;
    v.num_specs = ss->numSpecs;
    v.file_num = ss->fileNum;
    hcx_qbatch *qb = nullptr;
    status_t rc = hcx_queries_upload(ctx, &v, &qb);
    if (rc != SLM_SUCCESS) return rc;
    std::vector<hcx_result> res((size_t)ss->numSpecs);
    hcx_search_opts o = {};
    o.kernel = kernel;
    o.evalue = HCX_EVAL_TAILFIT;
    rc = hcx_search(ctx, ix.data(), (int)idxchunk, qb, &o, res.data(), nullptr, nullptr);
    hcx_queries_free(qb);
    if (rc != SLM_SUCCESS) return rc;
    for (int_t q = 0; q < ss->numSpecs; q++) {
        const hcx_result &r = res[q];
        if (!(r.flags & HCX_RES_VALID)) continue;      // cpsms < min_cpsm
        if (!(r.e_x < params.expect_max)) continue;
        hCell psm;
        psm.idxoffset = r.idxoffset;
        psm.sharedions = r.sharedions;
        psm.totalions = r.totalions;
        psm.fileIndex = (ushort_t)ss->fileNum;
        psm.psid = r.psid;
        psm.pmass = ss->precurse[q];
        psm.pchg = ss->charges ? ss->charges[q] : 0;
        psm.rtime = ss->rtimes ? ss->rtimes[q] : 0.0f;
        psm.hyperscore = r.hyperscore;
        rc = DFile_PrintScore(index, first + (uint_t)q, ss->precurse[q], &psm, r.e_x, r.cpsms);
        if (rc != SLM_SUCCESS) return rc;
    }
    return SLM_SUCCESS;
}

status_t DSLIM_ExportIonIndex(Index *index) {
    if (!index) return ERR_INVLD_PTR;
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.idx.find(index);
    if (it == s.idx.end()) return ERR_INVLD_PARAM;
    uint32_t peplen, speclen, n;
    uint64_t nions;
    status_t rc = hcx_index_info(it->second, &peplen, &speclen, &n, &nions);
    if (rc != SLM_SUCCESS) return rc;
    if (index->ionIndex) return ERR_INVLD_MEMORY;   // already holds host arrays
    SLMchunk *ch = new SLMchunk[1];
    ch->iA = new uint_t[nions ? nions : 1];
    ch->bA = new uint_t[(size_t)params.max_mass * params.scale + 1];
    rc = hcx_index_export(it->second, ch->iA, ch->bA);
    if (rc != SLM_SUCCESS) {
        delete[] ch->iA;
        delete[] ch->bA;
        delete[] ch;
        return rc;
    }
    index->ionIndex = ch;
    index->nChunks = 1;
    return SLM_SUCCESS;
}

// dslim.cpp:657-691: release one length's ion index (device, and the host
// copy if DSLIM_ExportIonIndex made one).
status_t DSLIM_DeallocateIonIndex(Index *index) {
    if (!index) return ERR_INVLD_PTR;
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.idx.find(index);
    if (it != s.idx.end()) {
        hcx_index_free(it->second);
        s.idx.erase(it);
    }
    if (index->ionIndex) {
        delete[] index->ionIndex[0].iA;
        delete[] index->ionIndex[0].bA;
        delete[] index->ionIndex;
        index->ionIndex = nullptr;
    }
    return SLM_SUCCESS;
}

// dslim.cpp:744-770: release the scorecard; here the whole device context,
// with every index still held.
status_t DSLIM_DeallocateSC() {
    State &s = st();
    std::lock_guard<std::mutex> g(s.mu);
    for (auto &kv : s.idx) hcx_index_free(kv.second);
    s.idx.clear();
    for (auto &kv : s.pdb) hcx_pepdb_free(kv.second.db);
    s.pdb.clear();
    if (s.ctx) hcx_destroy(s.ctx);
    s.ctx = nullptr;
    s.spectrumID = 0;
    return SLM_SUCCESS;
}

// expeRT.cpp:572-718: fit the survival-function tail of one spectrum's
// histogram (rPtr->survival holds counts) and store mu, beta, minhypscore
// and nexthypscore.  rPtr->cpsms and rPtr->maxhypscore are inputs.
status_t expeRT::ModelTailFit(Results *rPtr) {
    if (!rPtr || !rPtr->survival) return ERR_INVLD_PTR;
    State &s = st();
    hcx_ctx *ctx;
    uint_t hb;
    {
        std::lock_guard<std::mutex> g(s.mu);
        status_t rc = ensure_ctx();
        if (rc != SLM_SUCCESS) return rc;
        ctx = s.ctx;
        hb = 10 * s.max_hyperscore + 2;
    }
    std::vector<uint32_t> h(hb);
    for (uint_t i = 0; i < hb; i++) {
        const double v = rPtr->survival[i];
        if (!(v >= 0.0) || v > 4294967295.0 || v != (double)(uint32_t)v) return ERR_INVLD_PARAM;   // counts only
        h[i] = (uint32_t)v;
    }
    const uint32_t cps = rPtr->cpsms;
    const int32_t hyp = rPtr->maxhypscore;
    int32_t mu = 0, beta = 0, mn = 0, nx = 0;
    uint32_t fl = 0;
    status_t rc = hcx_tailfit_batch(ctx, h.data(), hb, 1, &cps, &hyp, &mu, &beta, &mn, &nx, &fl);
    if (rc != SLM_SUCCESS) return rc;
    rPtr->mu = mu;
    rPtr->beta = beta;
    rPtr->minhypscore = mn;
    rPtr->nexthypscore = nx;
    return SLM_SUCCESS;
}
