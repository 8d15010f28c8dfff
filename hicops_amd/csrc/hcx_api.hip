// hcx_api.hip -- the C ABI of libhcx (declared in include/hcx.h).
//
// Device memory is owned by the library: an index keeps its iA/bA/pepEntry
// arrays in HBM for its lifetime, a query batch keeps its spectra and result
// buffers.  Host buffers passed in are only read or written during the call.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "hcx_internal.h"

namespace hcx {

static thread_local std::string g_err;

int set_error(const std::string &msg) {
    g_err = msg;
    return HCX_ERR_DEVICE;
}
int set_error_code(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int DeviceBuf::ensure(size_t n) {
    if (n <= bytes && p) return 0;
    release();
    if (n == 0) n = 16;
    if (hipMalloc(&p, n) != hipSuccess) {
        p = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        return 1;
    }
    bytes = n;
    return 0;
}
void DeviceBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

size_t query_smem_bytes(const QueryArgs &a, bool replay);
int launch_candidates(hcx_ctx *ctx, const QueryArgs &a);

}  // namespace hcx

using namespace hcx;

extern "C" const char *hcx_last_error(void) { return g_err.c_str(); }

extern "C" int hcx_heap_capacity(uint32_t topmatches) {
    // dslim.cpp:422-428: 2^floor(log2(t) + 0.999) - 1
    uint32_t num = topmatches == 0 ? 1 : topmatches;
    num = (uint32_t)std::floor(std::log2((double)num) + 0.999);
    return (1 << num) - 1;
}

// ---- scratch of the allocation-free search (hcx_search_noalloc) ----
// A file-scope device array: every GPU the library runs on holds its own
// copy for the life of the process.  Its name is a host-side stand-in, not a
// device address: host code uses only the per-device pointer that
// hipGetSymbolAddress returns (scratch_ptr), never the name.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static void *g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

namespace {
// The device's scratch address, looked up once per device with that device
// current (hcx_create runs it).
int scratch_init(int device, void **ptr) {
    if (device < 0 || device >= kScratchDevices) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_create: device id");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(std::string("hcx_create: scratch address: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return HCX_SUCCESS;
}
}  // namespace

extern "C" int hcx_create(const hcx_params *params, int device, hcx_ctx **out) {
    if (!params || !out) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_create: null argument");
    const hcx_params &p = *params;
    if (p.maxz == 0 || p.maxz > 4 || p.scale == 0 || p.max_mass == 0 || p.max_mass <= p.min_mass)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_create: bad charge/scale/mass range");
    if ((uint64_t)p.max_mass * p.scale >= (1u << 30))
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_create: max_mass*scale too large");
    if (p.max_hyperscore == 0 || p.max_hyperscore > 1000 || p.max_shdpeaks == 0 || p.max_shdpeaks > 4096)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_create: bad max_hyperscore/max_shdpeaks");
    if (p.nmods > HCX_MAX_MOD_TYPES) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_create: too many mods");
    if ((uint64_t)p.max_mass * p.scale <= 2ull * p.dF + 2)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_create: dF too large");
    const int cap = hcx_heap_capacity(p.topmatches);
    if (cap <= 0 || cap > 255)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_create: topmatches gives an empty or huge heap");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        (void)hipGetLastError();
        return set_error_code(HCX_ERR_DEVICE, "hcx_create: no HIP device");
    }
    hcx_ctx *c = new (std::nothrow) hcx_ctx;
    if (!c) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_create: out of memory");
    c->params = p;
    c->device = device;
    c->NB = p.max_mass * p.scale;
    c->hbins = 10 * p.max_hyperscore + 2;
    c->heap_cap = cap;
    auto fail = [&](const char *what) {
        int rc = set_error(std::string("hcx_create: ") + what);
        hcx_destroy(c);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return fail("hipSetDevice");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail("hipGetDeviceProperties");
    c->num_cu = prop.multiProcessorCount;
    if (hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        c->max_lds = 65536;
    if (scratch_init(device, &c->scratch) != HCX_SUCCESS) {
        hcx_destroy(c);
        return HCX_ERR_DEVICE;
    }
    // one byte of hit counter per peptide of a chunk; the rest of a
    // workgroup's LDS share holds slots, peaks, run table, histogram and the
    // hit list (16384 measured best on a 160 KB CU with two workgroups each)
    c->chunk = c->max_lds >= 96 * 1024 ? 16384 : 8192;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail("stream");
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess ||
        hipEventCreate(&c->ev4) != hipSuccess || hipEventCreate(&c->ev5) != hipSuccess ||
        hipEventCreate(&c->ev6) != hipSuccess)
        return fail("events");
    // utils.h:115-131 lgfact table: val[i] = log10(i) + val[i-1]
    c->h_lgf.assign(p.max_shdpeaks + 1, 0.0);
    for (uint32_t i = 2; i <= p.max_shdpeaks; i++) c->h_lgf[i] = std::log10((double)i) + c->h_lgf[i - 1];
    if (hipMalloc(&c->d_lgf, sizeof(double) * c->h_lgf.size()) != hipSuccess) return fail("lgfact alloc");
    // staging for the host tail fits (fix_host_fits), allocated once here so
    // that no search allocates for it
    if (c->fix_dev.ensure((size_t)kFixRows * c->hbins * 4 + (size_t)kFixRows * sizeof(hcx_result)))
        return fail("host fit staging alloc");
    if (hipHostMalloc(&c->fix_host, fix_host_bytes(c->hbins), hipHostMallocDefault) != hipSuccess) {
        c->fix_host = nullptr;
        return fail("host fit staging alloc");
    }
    if (hipMemcpy(c->d_lgf, c->h_lgf.data(), sizeof(double) * c->h_lgf.size(), hipMemcpyHostToDevice) !=
        hipSuccess)
        return fail("lgfact copy");
    *out = c;
    return HCX_SUCCESS;
}

extern "C" int hcx_destroy(hcx_ctx *c) {
    if (!c) return HCX_SUCCESS;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->lens.release();
    c->tie.release();
    c->work.release();
    c->sur.release();
    c->sur_tmp.release();
    c->mlp_buf.release();
    c->fix_dev.release();
    if (c->fix_host) (void)hipHostFree(c->fix_host);
    if (c->d_lgf) (void)hipFree(c->d_lgf);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->ev3) (void)hipEventDestroy(c->ev3);
    if (c->ev4) (void)hipEventDestroy(c->ev4);
    if (c->ev5) (void)hipEventDestroy(c->ev5);
    if (c->ev6) (void)hipEventDestroy(c->ev6);
    for (hipStream_t a : c->aux)
        if (a) (void)hipStreamDestroy(a);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HCX_SUCCESS;
}

extern "C" int hcx_set_tuning(hcx_ctx *c, uint32_t chunk, uint32_t slots, uint32_t blocks_per_cu) {
    if (!c) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_set_tuning: null context");
    if (chunk) {
        if (chunk < 256 || chunk > 32768 || (chunk & (chunk - 1)))
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_tuning: chunk must be a power of two in [256, 32768]");
        c->chunk = chunk;
    }
    if (slots) {
        if (slots < 64 || slots > 32768) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_tuning: slots in [64, 32768]");
        c->slots = slots;
    }
    if (blocks_per_cu) {
        if (blocks_per_cu > 16) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_tuning: blocks_per_cu <= 16");
        c->blocks_per_cu = blocks_per_cu;
    }
    return HCX_SUCCESS;
}

extern "C" int hcx_set_option(hcx_ctx *c, int32_t option, int64_t value) {
    if (!c) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_set_option: null context");
    switch (option) {
    case HCX_OPT_HITLIST:
        c->hitlist = (int32_t)std::max<int64_t>(-1, std::min<int64_t>(value, 1 << 24));
        return HCX_SUCCESS;
    case HCX_OPT_SCAN:
        if (value != HCX_SCAN_AUTO && value != HCX_SCAN_TWO_PASS)
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_option: scan is HCX_SCAN_AUTO or HCX_SCAN_TWO_PASS");
        c->scan_mode = (int32_t)value;
        return HCX_SUCCESS;
    case HCX_OPT_GRAB:
        if (value < 1 || value > (int64_t)kMaxGrab) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_option: grab in 1..256");
        c->grab = (uint32_t)value;
        return HCX_SUCCESS;
    case HCX_OPT_FORCE_REDO:
        if (value < 0 || value > 0x7FFFFFFF) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_option: force_redo >= 0");
        c->force_redo = (uint32_t)value;
        return HCX_SUCCESS;
    case HCX_OPT_FORCE_HOST_FIT:
        if (value < 0 || value > 0x7FFFFFFF)
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_option: force_host_fit >= 0");
        c->force_host_fit = (uint32_t)value;
        return HCX_SUCCESS;
    case HCX_OPT_GROUP:
        if (value != 0 && (value < 64 || value > (int64_t)(kProbesPerThread * kChunkThreads)))
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_option: group is 0 (automatic) or 64..5 x the workgroup threads");
        c->group = (uint32_t)value;
        return HCX_SUCCESS;
    default:
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_set_option: unknown option");
    }
}

extern "C" int hcx_get_option(const hcx_ctx *c, int32_t option, int64_t *value) {
    if (!c || !value) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_get_option: null argument");
    switch (option) {
    case HCX_OPT_HITLIST: *value = c->hitlist; return HCX_SUCCESS;
    case HCX_OPT_SCAN: *value = c->scan_mode; return HCX_SUCCESS;
    case HCX_OPT_GRAB: *value = c->grab; return HCX_SUCCESS;
    case HCX_OPT_FORCE_REDO: *value = c->force_redo; return HCX_SUCCESS;
    case HCX_OPT_FORCE_HOST_FIT: *value = c->force_host_fit; return HCX_SUCCESS;
    case HCX_OPT_GROUP: *value = c->group; return HCX_SUCCESS;
    default: return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_get_option: unknown option");
    }
}

extern "C" int hcx_get_tuning(const hcx_ctx *c, uint32_t *chunk, uint32_t *slots, uint32_t *blocks_per_cu,
                              int32_t *num_cu, int32_t *max_lds) {
    if (!c) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_get_tuning: null context");
    if (chunk) *chunk = c->chunk;
    if (slots) *slots = c->slots;
    if (blocks_per_cu) *blocks_per_cu = c->blocks_per_cu;
    if (num_cu) *num_cu = c->num_cu;
    if (max_lds) *max_lds = c->max_lds;
    return HCX_SUCCESS;
}

extern "C" int hcx_lgfact_table(const hcx_ctx *c, double *out, uint32_t n) {
    if (!c || !out) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_lgfact_table: null");
    if (n > c->h_lgf.size()) return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_lgfact_table: n too large");
    std::memcpy(out, c->h_lgf.data(), sizeof(double) * n);
    return HCX_SUCCESS;
}

extern "C" int hcx_index_build(hcx_ctx *c, uint32_t peplen, const char *seqs, uint32_t nseqs,
                               const hcx_pep_entry *entries, uint32_t nentries, hcx_index **out) {
    if (!c || !out || (nentries && (!seqs || !entries)))
        return set_error_code(HCX_ERR_INVLD_PTR, "hcx_index_build: null argument");
    if (peplen < 2 || peplen > 63) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build: bad peptide length");
    const hcx_params &p = c->params;
    const uint32_t speclen = 2 * p.maxz * (peplen - 1);
    const uint64_t nions = (uint64_t)nentries * speclen;
    if (nions >= 0xFFFFFFF0ull)
        return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_index_build: more than 2^32 ions in one length");
    for (uint64_t i = 0; i < (uint64_t)nseqs * peplen; i++)
        if (seqs[i] < 'A' || seqs[i] > 'Z')
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build: sequence characters must be A-Z");
    for (uint32_t k = 0; k < nentries; k++) {
        const hcx_pep_entry &e = entries[k];
        if (e.seq_id >= nseqs)
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build: entry seq_id out of range");
        if (e.mod_num == 0) continue;
        // modAA (slm_dsts.h:120-144): one nibble (type + 1) per modified site
        int nib = 0;
        for (int i = 0; i < 8; i++) {
            uint32_t v = (e.mod_num >> (4 * i)) & 0xF;
            if (v == 0) break;
            if (v > p.nmods) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build: mod type out of range");
            nib++;
        }
        if ((nib < 8 && (e.mod_num >> (4 * nib)) != 0) || (e.sites >> peplen) != 0 ||
            __builtin_popcountll(e.sites) != nib)
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build: mod sites and mod_num disagree");
    }
    for (uint32_t k = 1; k < nentries; k++)
        if (!(entries[k - 1].mass <= entries[k].mass))
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build: entries not sorted by mass");
    if (hipSetDevice(c->device) != hipSuccess) return set_error("hipSetDevice");
    std::vector<float> mass(nentries);
    for (uint32_t k = 0; k < nentries; k++) mass[k] = entries[k].mass;
    return hcx::index_build_from(c, peplen, seqs, nseqs, entries, mass.data(), nentries, hipMemcpyHostToDevice, out);
}

namespace hcx {

// DSLIM_Construct over entries, masses and sequences copied from host or
// device memory (kind), into a new index.
int index_build_from(hcx_ctx *c, uint32_t peplen, const char *seqs, uint32_t nseqs, const hcx_pep_entry *entries,
                     const float *mass, uint32_t nentries, hipMemcpyKind kind, hcx_index **out) {
    const hcx_params &p = c->params;
    const uint32_t speclen = 2 * p.maxz * (peplen - 1);
    const uint64_t nions = (uint64_t)nentries * speclen;
    if (nions >= 0xFFFFFFF0ull)
        return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_index_build: more than 2^32 ions in one length");
    hcx_index *ix = new (std::nothrow) hcx_index;
    if (!ix) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_index_build: out of memory");
    ix->ctx = c;
    ix->peplen = peplen;
    ix->speclen = speclen;
    ix->n = nentries;
    ix->nseqs = nseqs;
    ix->nions = nions;
    const uint32_t D = p.max_mass + 2;
    if (ix->entries.ensure(sizeof(hcx_pep_entry) * (size_t)nentries + 16) ||
        ix->mass.ensure(4ull * nentries + 16) || ix->iA.ensure(4 * nions + 16) ||
        ix->bA.ensure(4ull * (c->NB + 1)) || ix->mdir.ensure(4ull * D) ||
        ix->seqs.ensure((size_t)nseqs * peplen + 16)) {
        delete ix;
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_index_build: device allocation failed");
    }
    auto fail = [&](int rc) {
        (void)hipStreamSynchronize(c->stream);
        delete ix;
        return rc;
    };
    if (nentries) {
        if (hipMemcpyAsync(ix->entries.p, entries, sizeof(hcx_pep_entry) * nentries, kind, c->stream) != hipSuccess ||
            hipMemcpyAsync(ix->mass.p, mass, 4ull * nentries, kind, c->stream) != hipSuccess ||
            hipMemcpyAsync(ix->seqs.p, seqs, (size_t)nseqs * peplen, kind, c->stream) != hipSuccess)
            return fail(set_error("hcx_index_build: upload failed"));
    }
    BuildArgs a;
    a.seqs = (const char *)ix->seqs.p;
    a.entries = (const hcx_pep_entry *)ix->entries.p;
    a.n = nentries;
    a.peplen = peplen;
    a.speclen = speclen;
    a.maxz = p.maxz;
    a.scale = p.scale;
    a.NB = c->NB;
    a.min_mass = (double)p.min_mass;
    a.max_mass = (double)p.max_mass;
    std::memcpy(a.mod_mass, p.mod_mass, sizeof(a.mod_mass));
    // every allocation before the timed device work: the (bin, chunk)
    // refinement, the chunk-local entries and the sort's scratch
    ix->chunk = c->chunk;
    ix->nch = std::max<uint32_t>(1, (nentries + c->chunk - 1) / c->chunk);
    if (ix->bA2.ensure(4ull * ba2_words(c->NB, ix->nch)))
        return fail(set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_index_build: chunk offsets allocation failed"));
    if (ix->iAc.ensure(2 * nions + 16))
        return fail(set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_index_build: chunk-local entry allocation failed"));
    if (int rr = reserve_index_build(c, nions, c->NB)) return fail(rr);
    if (hipEventRecord(c->ev2, c->stream) != hipSuccess) return fail(set_error("event"));
    int rc = launch_index_build(c, a, (uint32_t *)ix->iA.p, (uint32_t *)ix->bA.p, (const float *)ix->mass.p,
                                (uint32_t *)ix->mdir.p, p.max_mass);
    if (rc) return fail(rc);
    // (bin, chunk) refinement of bA for the chunked query
    rc = launch_chunk_offsets(c, (const uint32_t *)ix->iA.p, (const uint32_t *)ix->bA.p, c->NB, ix->nch, ix->chunk,
                              speclen, (uint32_t *)ix->bA2.p);
    if (rc) return fail(rc);
    rc = launch_chunk_local(c, (const uint32_t *)ix->iA.p, nions, speclen, ix->chunk, (uint16_t *)ix->iAc.p);
    if (rc) return fail(rc);
    if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess)
        return fail(set_error("event"));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev2, c->ev1);
    ix->build_ms = ms;
    *out = ix;
    return HCX_SUCCESS;
}

int index_build_device(hcx_ctx *c, uint32_t peplen, const char *d_seqs, uint32_t nseqs, const hcx_pep_entry *d_entries,
                       const float *d_mass, uint32_t nentries, hcx_index **out) {
    if (hipSetDevice(c->device) != hipSuccess) return set_error("hipSetDevice");
    return index_build_from(c, peplen, d_seqs, nseqs, d_entries, d_mass, nentries, hipMemcpyDeviceToDevice, out);
}

}  // namespace hcx

extern "C" int hcx_index_info(const hcx_index *ix, uint32_t *peplen, uint32_t *speclen, uint32_t *nentries,
                              uint64_t *nions) {
    if (!ix) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_index_info: null");
    if (peplen) *peplen = ix->peplen;
    if (speclen) *speclen = ix->speclen;
    if (nentries) *nentries = ix->n;
    if (nions) *nions = ix->nions;
    return HCX_SUCCESS;
}

extern "C" int hcx_index_build_ms(const hcx_index *ix, float *ms) {
    if (!ix || !ms) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_index_build_ms: null");
    *ms = ix->build_ms;
    return HCX_SUCCESS;
}

extern "C" int hcx_index_export(const hcx_index *ix, uint32_t *iA, uint32_t *bA) {
    if (!ix) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_index_export: null");
    hcx_ctx *c = ix->ctx;
    HCX_HIP(hipSetDevice(c->device));
    HCX_HIP(hipStreamSynchronize(c->stream));
    if (iA && ix->nions) HCX_HIP(hipMemcpy(iA, ix->iA.p, 4 * ix->nions, hipMemcpyDeviceToHost));
    if (bA) HCX_HIP(hipMemcpy(bA, ix->bA.p, 4ull * (c->NB + 1), hipMemcpyDeviceToHost));
    return HCX_SUCCESS;
}

extern "C" int hcx_index_free(hcx_index *ix) {
    if (!ix) return HCX_SUCCESS;
    (void)hipSetDevice(ix->ctx->device);
    (void)hipStreamSynchronize(ix->ctx->stream);
    delete ix;
    return HCX_SUCCESS;
}

extern "C" int hcx_queries_upload(hcx_ctx *c, const hcx_queries_view *q, hcx_qbatch **out) {
    if (!c || !q || !out) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_queries_upload: null argument");
    if (q->num_specs < 0) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_queries_upload: negative count");
    const int32_t n = q->num_specs;
    if (n > 0 && (!q->idx || !q->precurse || !q->charges))
        return set_error_code(HCX_ERR_INVLD_PTR, "hcx_queries_upload: missing arrays");
    const uint64_t npk = n > 0 ? q->idx[n] : 0;
    if (npk && (!q->moz || !q->intensity))
        return set_error_code(HCX_ERR_INVLD_PTR, "hcx_queries_upload: missing peaks");
    uint32_t maxp = 0;
    for (int32_t s = 0; s < n; s++) {
        if (q->idx[s + 1] < q->idx[s]) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_queries_upload: idx not monotone");
        maxp = std::max(maxp, q->idx[s + 1] - q->idx[s]);
    }
    if (n > 0 && q->idx[0] != 0) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_queries_upload: idx[0] != 0");
    HCX_HIP(hipSetDevice(c->device));
    hcx_qbatch *b = new (std::nothrow) hcx_qbatch;
    if (!b) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_queries_upload: out of memory");
    b->ctx = c;
    b->nspec = n;
    b->file_num = q->file_num;
    b->npeaks = npk;
    b->max_peaks = maxp;
    if (b->moz.ensure(4 * npk + 16) || b->inten.ensure(4 * npk + 16) || b->idx.ensure(4ull * (n + 1)) ||
        b->prec.ensure(4ull * n + 16) || b->chg.ensure(4ull * n + 16) || b->rt.ensure(4ull * n + 16)) {
        delete b;
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_queries_upload: device allocation failed");
    }
    auto up = [&](hcx::DeviceBuf &d, const void *h, size_t bytes) {
        if (!bytes || !h) return hipSuccess;
        return hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c->stream);
    };
    // a batch without retention times holds zeros (what hcx_queries_export returns)
    const bool zero_rt = !q->rtimes && n > 0;
    if (up(b->moz, q->moz, 4 * npk) || up(b->inten, q->intensity, 4 * npk) || up(b->idx, q->idx, 4ull * (n + 1)) ||
        up(b->prec, q->precurse, 4ull * n) || up(b->chg, q->charges, 4ull * n) || up(b->rt, q->rtimes, 4ull * n) ||
        (zero_rt && hipMemsetAsync(b->rt.p, 0, 4ull * n, c->stream) != hipSuccess)) {
        delete b;
        return set_error("hcx_queries_upload: copy failed");
    }
    // largest intensity and, per spectrum, the most peaks and the largest
    // intensity sum inside any 2*dF+1 bins (they bound what one fragment bin
    // can collect), on the device
    if (int rc = hcx::batch_facts(c, b, maxp)) {
        delete b;
        return rc;
    }
    *out = b;
    return HCX_SUCCESS;
}

extern "C" int hcx_queries_free(hcx_qbatch *b) {
    if (!b) return HCX_SUCCESS;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return HCX_SUCCESS;
}

namespace {

// Fill QueryArgs for a search; also sizes the batch's result buffers.
// Device memory one search uses beyond its inputs: its outputs (results,
// histograms, work counters) and its scratch.  hcx_search and its variants
// take it from grow-only buffers of the context and the batch;
// hcx_search_noalloc carves it from the static scratch array.
struct SearchMem {
    LenDesc *lens;
    hcx_result *res;
    uint32_t *hist;
    hcx_stats *stats;
    uint32_t *tie, *queue, *scan_tmp;
    SpecAcc *acc;
    int2 *win;
    uint8_t *packed;
    uint32_t *fb, *counts, *item_base, *item_spec, *item_max;
    uint32_t item_cap;
    uint32_t *redo, *tie2;   // spectra re-scored on the cursor kernel, and their ties
};

// the item_spec entries a search of n spectra over indices of nch_all chunks needs
uint64_t item_spec_words(uint64_t n, uint64_t nch_all) { return (n * nch_all) / 64 + 2; }

int search_mem_alloc(hcx_ctx *c, hcx_qbatch *qb, int nidx, uint64_t nch_all, int kernel, SearchMem *m) {
    const uint64_t n = (uint64_t)qb->nspec;
    if (c->lens.ensure(sizeof(LenDesc) * std::max(nidx, 1)) || qb->res.ensure(sizeof(hcx_result) * n + 16) ||
        qb->hist.ensure(4ull * c->hbins * n + 16) || qb->stats.ensure(sizeof(hcx_stats) * n + 16) ||
        c->tie.ensure(4 * (n + 2) + 16) || c->work.ensure(64 + 4ull * kSurSlots * kSurStride) || qb->acc.ensure(sizeof(SpecAcc) * n + 16) ||
        qb->win.ensure(8ull * n * std::max(nidx, 1) + 16) || qb->packed.ensure(n + 16) ||
        qb->fb.ensure(4 * (n + 1) + 16) || qb->counts.ensure(4 * (n + 1) + 16) ||
        qb->item_base.ensure(4 * (n + 1) + 16) || qb->item_spec.ensure(4 * item_spec_words(n, nch_all) + 16) ||
        c->scan_tmp.ensure(4 * scan_tmp_words(n + 1) + 64) || qb->redo.ensure(8 * (n + 2) + 16))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "search: device allocation failed");
    if (kernel == HCX_KERNEL_BATCHED && !qb->item_max.p) {
        // a first estimate; grown to the item total after each search
        qb->item_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(64 * n, 1024), 0xFFFFFFF0ull);
        if (qb->item_max.ensure(4ull * qb->item_cap + 16))
            return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "search: device allocation failed");
    }
    m->lens = (LenDesc *)c->lens.p;
    m->res = (hcx_result *)qb->res.p;
    m->hist = (uint32_t *)qb->hist.p;
    m->stats = (hcx_stats *)qb->stats.p;
    m->tie = (uint32_t *)c->tie.p;
    m->queue = (uint32_t *)c->work.p;
    m->scan_tmp = (uint32_t *)c->scan_tmp.p;
    m->acc = (SpecAcc *)qb->acc.p;
    m->win = (int2 *)qb->win.p;
    m->packed = (uint8_t *)qb->packed.p;
    m->fb = (uint32_t *)qb->fb.p;
    m->counts = (uint32_t *)qb->counts.p;
    m->item_base = (uint32_t *)qb->item_base.p;
    m->item_spec = (uint32_t *)qb->item_spec.p;
    m->item_max = (uint32_t *)qb->item_max.p;
    m->item_cap = qb->item_max.p ? qb->item_cap : 0;
    m->redo = (uint32_t *)qb->redo.p;
    m->tie2 = m->redo + (n + 2);
    return HCX_SUCCESS;
}

// The scratch of a search of n spectra over nidx indices (nch_all chunks)
// carved from [base, base + bytes); item_max takes what is left.  *need: the
// bytes required before item_max.
bool search_mem_carve(void *base, size_t bytes, uint64_t n, int nidx, uint64_t nch_all, SearchMem *m, size_t *need) {
    size_t off = 0;
    auto take = [&](size_t sz) {
        const size_t o = off;
        off = (off + sz + 255) & ~(size_t)255;
        return (unsigned char *)base + o;
    };
    m->lens = (LenDesc *)take(sizeof(LenDesc) * std::max(nidx, 1));
    m->tie = (uint32_t *)take(4 * (n + 2));
    m->queue = (uint32_t *)take(64);
    m->scan_tmp = (uint32_t *)take(4 * scan_tmp_words(n + 1) + 64);
    m->acc = (SpecAcc *)take(sizeof(SpecAcc) * n);
    m->win = (int2 *)take(8 * n * std::max(nidx, 1));
    m->packed = (uint8_t *)take(n);
    m->fb = (uint32_t *)take(4 * (n + 1));
    m->counts = (uint32_t *)take(4 * (n + 1));
    m->item_base = (uint32_t *)take(4 * (n + 1));
    m->item_spec = (uint32_t *)take(4 * item_spec_words(n, nch_all));
    m->redo = (uint32_t *)take(4 * (n + 2));
    m->tie2 = (uint32_t *)take(4 * (n + 2));
    *need = off;
    if (off >= bytes) return false;
    m->item_max = (uint32_t *)((unsigned char *)base + off);
    m->item_cap = (uint32_t)std::min<uint64_t>((bytes - off) / 4, 0xFFFFFFF0ull);
    return true;
}

int prepare(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb, const hcx_search_opts *opts,
            QueryArgs *a, const SearchMem &m) {
    if (!c || !qb || (nidx > 0 && !idx)) return set_error_code(HCX_ERR_INVLD_PTR, "search: null argument");
    if (qb->ctx != c) return set_error_code(HCX_ERR_INVLD_PARAM, "search: batch from another context");
    if (nidx < 0 || nidx > 64) return set_error_code(HCX_ERR_INVLD_PARAM, "search: 0..64 indices");
    const hcx_params &p = c->params;
    LenDesc lens[64];
    const int nlens = nidx > 0 ? nidx : 1;
    std::memset(lens, 0, sizeof lens);
    uint32_t nbk = 0;
    for (int i = 0; i < nidx; i++) {
        const hcx_index *ix = idx[i];
        if (!ix || ix->ctx != c) return set_error_code(HCX_ERR_INVLD_PARAM, "search: bad index");
        LenDesc &L = lens[i];
        L.mass = (const float *)ix->mass.p;
        L.iA = (const uint32_t *)ix->iA.p;
        L.bA = (const uint32_t *)ix->bA.p;
        L.mdir = (const uint32_t *)ix->mdir.p;
        L.n = ix->n;
        L.speclen = ix->speclen;
        L.half = ix->speclen / 2;
        L.peplen = ix->peplen;
        make_divider(ix->speclen, &L.div_m, &L.div_s);
        L.nch = ix->nch;
        L.cbase = nbk;
        nbk += ix->nch;
        if (ix->nch > 0xFFFF || nbk >= (1u << 30))
            return set_error_code(HCX_ERR_INVLD_SIZE, "search: too many peptide chunks for the chunk size");
        L.bA2 = (const uint32_t *)ix->bA2.p;
        L.iAc = (const uint16_t *)ix->iAc.p;
        if (ix->chunk != c->chunk)
            return set_error_code(HCX_ERR_INVLD_PARAM, "search: index built with another chunk size");
        if (ix->n >= (1u << 25)) return set_error_code(HCX_ERR_INVLD_SIZE, "search: more than 2^25 entries in a length");
        // the scan reads iAc through 32-bit byte offsets
        if ((uint64_t)ix->n * ix->speclen * 2 > 0xFFFFFFFFull)
            return set_error_code(HCX_ERR_INVLD_SIZE, "search: a length's ion array exceeds 4 GiB");
    }
    HCX_HIP(hipMemcpyAsync(m.lens, lens, sizeof(LenDesc) * nlens, hipMemcpyHostToDevice, c->stream));
    std::memset(a, 0, sizeof(*a));
    a->lens = m.lens;
    a->nlen = nidx;
    a->moz = (const uint32_t *)qb->moz.p;
    a->inten = (const uint32_t *)qb->inten.p;
    a->qidx = (const uint32_t *)qb->idx.p;
    a->prec = (const float *)qb->prec.p;
    a->nspec = qb->nspec;
    a->dF = p.dF;
    a->NB = c->NB;
    a->min_shp = p.min_shp;
    a->hbins = c->hbins;
    a->max_shdpeaks = p.max_shdpeaks;
    a->max_mass = p.max_mass;
    a->dM = p.dM;
    a->lgf = c->d_lgf;
    a->res = m.res;
    a->hist = m.hist;
    a->stats = m.stats;
    a->tie = m.tie;
    a->heap_cap = c->heap_cap;
    a->nprobe_max = std::max<uint32_t>(1, qb->max_peaks) * (2 * p.dF + 1);
    // tile: the largest power of two (<= 8192) whose LDS layout fits 80 KB,
    // or the caller's choice if it fits the device limit
    int dev_lds = 65536;
    (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    uint32_t tile = opts && opts->tile ? opts->tile : 0;
    if (tile) {
        a->tile = tile;
        if (query_smem_bytes(*a, true) > (size_t)dev_lds)
            return set_error_code(HCX_ERR_INVLD_PARAM, "search: tile does not fit in LDS");
    } else {
        const size_t budget = std::min<size_t>((size_t)dev_lds, 80 * 1024);
        for (tile = 8192; tile >= 64; tile /= 2) {
            a->tile = tile;
            if (query_smem_bytes(*a, true) <= budget) break;
        }
        if (tile < 64) return set_error_code(HCX_ERR_INVLD_PARAM, "search: spectra too long for LDS");
    }
    return HCX_SUCCESS;
}

uint64_t fnv1a(const void *p, size_t n, uint64_t h) {
    const unsigned char *b = (const unsigned char *)p;
    h ^= 1469598103934665603ull;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// One search's scan plan: the two-pass kernel's arguments (ca), the packed
// kernel's (pa: the same apart from its run-table group), the window kernel's
// outputs, and which kernels run.
struct ScanPlan {
    ChunkArgs ca, pa;
    WinArgs wa;
    int kernel = HCX_KERNEL_PLAIN;
    bool pack = false;       // packed kernel for the spectra it scores exactly
    uint32_t *scan_tmp = nullptr;
    uint32_t *tie2 = nullptr;   // ties of the spectra re-scored on the cursor kernel
};

// The device work of one search, all on the context stream with no host
// synchronisation: zero the per-spectrum state with fill kernels (a memset
// node is not used, so the sequence can be captured into a graph), windows
// and work items, scan, replay tied spectra, tail fit.  ev0 / ev6 / ev1
// bracket the packed and the two-pass scan kernels.
int launch_sequence(hcx_ctx *c, ScanPlan &sp, bool tailfit = true) {
    ChunkArgs &ca = sp.ca;
    QueryArgs &a = ca.q;
    const uint64_t n = (uint64_t)a.nspec;
    const bool batched = sp.kernel == HCX_KERNEL_BATCHED;
    int rc = launch_fill_u32(c, (uint32_t *)ca.acc, n * (sizeof(SpecAcc) / 4), 0);
    if (!rc) rc = launch_fill_u32(c, a.hist, n * c->hbins, 0);
    if (!rc) rc = launch_fill_u32(c, a.tie, 1, 0);
    // the replay's work counter, after the list's nspec entries
    if (!rc) rc = launch_fill_u32(c, a.tie + a.nspec + 1, 1, 0);
    if (!rc) rc = launch_fill_u32(c, ca.queue, 1, 0);
    if (rc) return rc;
    if (sp.kernel == HCX_KERNEL_CURSOR) {
        if (!c->capturing) HCX_HIP(hipEventRecord(c->ev0, c->stream));
        if (!c->capturing) HCX_HIP(hipEventRecord(c->ev6, c->stream));
        rc = launch_cursor(c, a);          // writes results + tie list itself
        if (!rc && !c->capturing) HCX_HIP(hipEventRecord(c->ev1, c->stream));
        if (!rc) rc = launch_replay_cursor(c, a);   // its own (independent) replay
    } else {
        rc = launch_fill_u32(c, sp.wa.fb_count, 1, 0);
        if (!rc) rc = launch_fill_u32(c, ca.redo, 1, 0);
        if (!rc) rc = launch_fill_u32(c, sp.tie2, 1, 0);
        if (!rc && batched) rc = launch_fill_u32(c, sp.wa.counts + n, 1, 0);
        if (!rc) rc = launch_windows(c, ca, sp.wa);
        if (!rc && batched) {
            rc = scan_exclusive_u32(c->stream, sp.wa.counts, (uint32_t *)ca.item_base, n + 1, sp.scan_tmp);
            if (!rc) rc = launch_item_spec(c, ca.item_base, a.nspec, (uint32_t *)ca.item_spec);
        }
        if (rc) return rc;
        if (!c->capturing) HCX_HIP(hipEventRecord(c->ev0, c->stream));
        if (sp.pack) rc = launch_packed(c, sp.pa, sp.kernel);
        if (!rc && !c->capturing) HCX_HIP(hipEventRecord(c->ev6, c->stream));
        if (!rc) {
            if (sp.pack) {
                // the rest, one spectrum at a time from the list
                ChunkArgs fa = ca;
                fa.fb_list = sp.wa.fb_list;
                fa.fb_count = sp.wa.fb_count;
                rc = launch_chunked(c, fa, HCX_KERNEL_PLAIN);
            } else {
                rc = launch_chunked(c, ca, sp.kernel);
            }
        }
        if (!rc && !c->capturing) HCX_HIP(hipEventRecord(c->ev1, c->stream));
        if (!rc) rc = launch_finalize(c, ca);
        if (!rc) rc = launch_replay_chunked(c, ca);
        // spectra whose two-pass counters neared overflow (k_finalize lists
        // them): the cursor kernel scores them exactly and replays their ties
        if (!rc) {
            QueryArgs ra = a;
            ra.redo = ca.redo;
            ra.tie = sp.tie2;
            rc = launch_cursor_redo(c, ra);
            if (!rc) rc = launch_replay_cursor(c, ra);
        }
    }
    if (rc || !tailfit) return rc;
    if (!c->capturing) HCX_HIP(hipEventRecord(c->ev4, c->stream));
    // the redo list is done with: it now lists the fits for the host
    rc = launch_fill_u32(c, ca.redo, 1, 0);
    if (!rc)
        rc = launch_tailfit(c, a.res, a.hist, a.hbins, a.nspec, c->params.min_cpsm, c->params.expect_max, ca.redo,
                            c->force_host_fit);
    if (rc) return rc;
    if (!c->capturing) HCX_HIP(hipEventRecord(c->ev5, c->stream));
    return HCX_SUCCESS;
}

// Fill sp for a search of qb over idx (prepare() done): layouts, the scan
// plan and every device buffer the sequence uses.
int plan_search(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb, int kernel, const SearchMem &m,
                ScanPlan &sp) {
    ChunkArgs &ca = sp.ca;
    if (c->params.min_shp > kMaxChunkShp) kernel = HCX_KERNEL_CURSOR;
    sp.kernel = kernel;
    const uint64_t n = (uint64_t)qb->nspec;
    ca.tc = c->chunk;
    ca.slots = c->slots;
    ca.max_peaks = std::max<uint32_t>(1, qb->max_peaks);
    ca.grab = c->grab;
    make_divider(2 * c->params.dF + 1, &ca.wdiv_m, &ca.wdiv_s);
    uint32_t maxspec = 0;
    for (int i = 0; i < nidx; i++) maxspec = std::max(maxspec, idx[i]->speclen);
    // 32-bit slot sums when no peptide's sum can reach 2^32: at most
    // speclen * max_mult hits per peptide, each at most the batch's largest
    // intensity
    {
        const uint64_t hits = (uint64_t)maxspec * std::max<uint32_t>(1, qb->max_mult);
        ca.wide = (uint64_t)qb->max_int * hits >= (1ull << 32) ? 1u : 0u;
    }
    ca.hl_cap = 0;
    const uint32_t need = std::max<uint32_t>(1, ca.max_peaks * (2 * c->params.dF + 1));
    {
        const size_t share = (size_t)c->max_lds / std::max<uint32_t>(1, c->blocks_per_cu);
        // run-table group: an item walks its probes group by group, and every
        // extra group costs the exact-count pass another run setup, so by
        // default the group is as large as the workgroup's LDS share allows
        // (12 bytes per probe), but not below 1024 probes
        if (c->group) {
            ca.group = std::min<uint32_t>(c->group, need);
        } else {
            const uint32_t floor = std::min<uint32_t>(need, 1024);
            ca.group = std::min<uint32_t>(need, kProbesPerThread * kChunkThreads);
            while (ca.group > floor && chunk_smem_bytes(ca) > share) ca.group = std::max<uint32_t>(floor, ca.group - 128);
        }
        // the hit list (16-bit entries) takes the LDS left in a workgroup's
        // share of the CU once blocks_per_cu workgroups are resident
        const size_t base = chunk_smem_bytes(ca);
        if (share > base + 4096) ca.hl_cap = (uint32_t)(((share - base) / 2) & ~(size_t)63);
        if (c->hitlist >= 0) ca.hl_cap = std::min<uint32_t>(ca.hl_cap, (uint32_t)c->hitlist & ~63u);
    }
    ca.grid = (uint32_t)std::max(1, c->num_cu) * c->blocks_per_cu;
    if (chunk_smem_bytes(ca) > (size_t)c->max_lds)
        return set_error_code(HCX_ERR_INVLD_PARAM, "search: chunk/slots do not fit in LDS");
    // hits per peptide must fit the 15-bit counters: speclen * peaks within +-dF
    if ((uint64_t)maxspec * qb->max_mult >= 0x8000)
        return set_error_code(HCX_ERR_INVLD_PARAM, "search: too many peaks within +-dF of one bin");
    ca.acc = m.acc;
    ca.queue = m.queue;
    ca.redo = m.redo;
    ca.force_redo = c->force_redo;
    sp.tie2 = m.tie2;
    if (kernel == HCX_KERNEL_CURSOR) return HCX_SUCCESS;

    // the packed kernel: the largest run-table group that fits beside the
    // accumulators (one 1024-thread workgroup per CU)
    // (each spectrum is checked against the packed words' exactness bound in
    // k_windows; the rest go to the two-pass kernel)
    sp.pack = c->scan_mode == HCX_SCAN_AUTO;
    if (sp.pack) {
        ChunkArgs &pa = sp.pa;
        pa = ca;
        // (a run table larger than the batch needs is not wasted: the scan's
        // candidate queues use its space)
        pa.group = 2 * kPackedThreads;
        const size_t pshare = (size_t)c->max_lds / kPackedBlocks;
        while (pa.group > 64 && (!packed_smem_bytes(pa) || packed_smem_bytes(pa) > pshare)) pa.group -= 64;
        sp.pack = packed_smem_bytes(pa) && packed_smem_bytes(pa) <= pshare;
        pa.grid = (uint32_t)std::max(1, c->num_cu) * kPackedBlocks;
    }
    // windows, work-item numbering, the two-pass list
    if (!qb->mult.p || !qb->wsum.p)
        return set_error_code(HCX_ERR_INVLD_PARAM, "search: batch without peak multiplicities");
    WinArgs &wa = sp.wa;
    std::memset(&wa, 0, sizeof wa);
    wa.mult = (const uint32_t *)qb->mult.p;
    wa.wsum = (const uint32_t *)qb->wsum.p;
    wa.half = maxspec / 2;
    wa.pack = sp.pack ? 1u : 0u;
    wa.batched = kernel == HCX_KERNEL_BATCHED ? 1u : 0u;
    wa.win = m.win;
    wa.counts = m.counts;
    wa.packed = m.packed;
    wa.fb_list = m.fb + 1;
    wa.fb_count = m.fb;
    wa.acc = ca.acc;
    ca.win = wa.win;
    ca.packed = wa.packed;
    if (kernel == HCX_KERNEL_BATCHED) {
        ca.item_base = m.item_base;
        ca.item_spec = m.item_spec;
        ca.item_max = m.item_max;
        ca.item_cap = m.item_cap;
    }
    sp.scan_tmp = m.scan_tmp;
    ChunkArgs &pa = sp.pa;
    const uint32_t pgroup = pa.group, pgrid = pa.grid;
    if (sp.pack) {
        pa = ca;
        pa.group = pgroup;
        pa.grid = pgrid;
    }
    return HCX_SUCCESS;
}

// After a batched search: grow item_max to the search's item total, so that
// the next search of the batch records (and the replay can skip) every item.
int grow_item_max(hcx_ctx *c, hcx_qbatch *qb, const ScanPlan &sp) {
    if (sp.kernel != HCX_KERNEL_BATCHED) return HCX_SUCCESS;
    uint32_t total = 0;
    HCX_HIP(hipMemcpyAsync(&total, (const uint32_t *)qb->item_base.p + qb->nspec, 4, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    if (total > qb->item_cap) {
        qb->item_max.release();
        if (qb->item_max.ensure(4ull * total + 16)) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "search: alloc");
        qb->item_cap = total;
    }
    return HCX_SUCCESS;
}

// The indices' total chunk count, after checking the arguments a search
// needs before its memory is sized.
int search_shape(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb, uint64_t *nch_all) {
    if (!c || !qb || (nidx > 0 && !idx)) return set_error_code(HCX_ERR_INVLD_PTR, "search: null argument");
    if (nidx < 0 || nidx > 64) return set_error_code(HCX_ERR_INVLD_PARAM, "search: 0..64 indices");
    uint64_t t = 0;
    for (int i = 0; i < nidx; i++) {
        if (!idx[i]) return set_error_code(HCX_ERR_INVLD_PARAM, "search: bad index");
        t += idx[i]->nch;
    }
    *nch_all = t;
    return HCX_SUCCESS;
}

// One search.  mem: the device memory to use (hcx_search_noalloc), or null:
// the context's and the batch's own buffers.
int run_search(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb, const hcx_search_opts *opts,
               bool timed, float *kernel_ms, float *total_ms, const SearchMem *mem = nullptr) {
    uint64_t nch_all = 0;
    int rc = search_shape(c, idx, nidx, qb, &nch_all);
    if (rc) return rc;
    HCX_HIP(hipSetDevice(c->device));
    const int kernel0 = opts ? opts->kernel : HCX_KERNEL_PLAIN;
    SearchMem m;
    if (mem) {
        m = *mem;
    } else {
        rc = search_mem_alloc(c, qb, nidx, nch_all, kernel0, &m);
        if (rc) return rc;
    }
    ScanPlan sp;
    std::memset(&sp.ca, 0, sizeof sp.ca);
    std::memset(&sp.pa, 0, sizeof sp.pa);
    rc = prepare(c, idx, nidx, qb, opts, &sp.ca.q, m);
    if (rc) return rc;
    QueryArgs &a = sp.ca.q;
    const bool sur = opts && opts->evalue == HCX_EVAL_SURROGATE;
    if (opts && opts->evalue != HCX_EVAL_TAILFIT && !sur)
        return set_error_code(HCX_ERR_INVLD_PARAM, "search: unknown e-value method");
    if (sur && !c->sur_ready) return set_error_code(HCX_ERR_INVLD_PARAM, "search: the surrogate is not trained");
    rc = plan_search(c, idx, nidx, qb, kernel0, m, sp);
    if (rc) return rc;
    const int kernel = sp.kernel;
    ChunkArgs &ca = sp.ca;
    if (timed) HCX_HIP(hipEventRecord(c->ev2, c->stream));
    if (sur) {
        // scan, then predicted e-values, the audit sample's plain fits, and
        // the plain fit for everything if the audit finds a disagreement
        rc = launch_sequence(c, sp, false);
        // audit counters (4 words); spectra located / plain-fitted, counted
        // in kSurSlots slots a cache line apart
        uint32_t *cnt = (uint32_t *)c->work.p + 4;
        uint32_t *scnt = (uint32_t *)c->work.p + 16;
        if (!rc) rc = launch_fill_u32(c, cnt, 4, 0);
        if (!rc) rc = launch_fill_u32(c, scnt, kSurSlots * kSurStride, 0);
        if (!rc) rc = launch_fill_u32(c, ca.redo, 1, 0);
        if (!rc) HCX_HIP(hipEventRecord(c->ev4, c->stream));
        if (!rc) rc = mlp_infer(c, qb, ca.redo, scnt);
        if (!rc) HCX_HIP(hipEventRecord(c->ev5, c->stream));
        const uint32_t every = opts->audit_every ? opts->audit_every : 64;
        const double tol = opts->audit_tol > 0.0 ? opts->audit_tol : 1e-3;
        if (!rc)
            rc = launch_tailfit_audit(c, a.res, a.hist, a.hbins, a.nspec, c->params.min_cpsm, c->params.expect_max,
                                      every, tol, cnt, ca.redo, c->force_host_fit);
        if (rc) return rc;
        uint32_t h[12 + kSurSlots * kSurStride];   // cnt (work word 4) through the slots (word 16)
        HCX_HIP(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HCX_HIP(hipStreamSynchronize(c->stream));
        hcx_audit au = {};
        au.audited = h[0];
        au.disagreed = h[1];
        for (uint32_t k = 0; k < kSurSlots; k++) {
            au.located += h[12 + k * kSurStride];
            au.plain_fit += h[12 + k * kSurStride + 1];
        }
        uint64_t eb = (uint64_t)h[2] | ((uint64_t)h[3] << 32);
        std::memcpy(&au.max_rel_err, &eb, 8);
        au.fell_back = h[1] > 0 ? 1 : 0;
        c->last_audit = au;
        if (au.fell_back) {
            rc = launch_fill_u32(c, ca.redo, 1, 0);
            if (!rc)
                rc = launch_tailfit(c, a.res, a.hist, a.hbins, a.nspec, c->params.min_cpsm, c->params.expect_max,
                                    ca.redo, c->force_host_fit);
            if (rc) return rc;
        }
    } else if (opts && opts->use_graph) {
        // one graph per launch shape: every pointer and size the kernels see
        // is in the plan, so its bytes (zero-padded above) key the captured
        // graph
        uint64_t key = fnv1a(&ca, sizeof ca, (uint64_t)kernel * 0x9E3779B97F4A7C15ull + 1);
        key = fnv1a(&sp.pa, sizeof sp.pa, key);
        key = fnv1a(&sp.wa, sizeof sp.wa, key + (sp.pack ? 1 : 0));
        if (!c->graph_exec || c->graph_key != key) {
            if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
            c->graph_exec = nullptr;
            c->graph_key = 0;
            HCX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            c->capturing = true;   // no event records inside the graph
            rc = launch_sequence(c, sp);
            c->capturing = false;
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return set_error(std::string("search: stream capture: ") + hipGetErrorString(ec));
            hipGraphExec_t ex = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return set_error(std::string("search: graph instantiate: ") + hipGetErrorString(ei));
            c->graph_exec = ex;
            c->graph_key = key;
        }
        // a replayed graph is timed as a whole: scan and e-value times both
        // cover the captured sequence
        HCX_HIP(hipEventRecord(c->ev0, c->stream));
        HCX_HIP(hipEventRecord(c->ev4, c->stream));
        HCX_HIP(hipGraphLaunch(c->graph_exec, c->stream));
        HCX_HIP(hipEventRecord(c->ev6, c->stream));
        HCX_HIP(hipEventRecord(c->ev1, c->stream));
        HCX_HIP(hipEventRecord(c->ev5, c->stream));
    } else {
        rc = launch_sequence(c, sp);
        if (rc) return rc;
    }
    // fits the device could not decide: on the host, before the search ends
    rc = fix_host_fits(c, ca.redo, a.res, a.hist, a.hbins, c->params.expect_max);
    if (rc) return rc;
    if (timed) {
        HCX_HIP(hipEventRecord(c->ev3, c->stream));
        HCX_HIP(hipEventSynchronize(c->ev3));
        float q = 0.0f, t = 0.0f;
        HCX_HIP(hipEventElapsedTime(&q, c->ev0, c->ev1));
        HCX_HIP(hipEventElapsedTime(&t, c->ev2, c->ev3));
        if (kernel_ms) *kernel_ms = q;
        if (total_ms) *total_ms = t;
        c->timed_ok = true;
        c->ran_packed = sp.pack;
        c->ran_two_pass = true;   // the two-pass kernel (or the cursor kernel) is always launched
    } else {
        HCX_HIP(hipStreamSynchronize(c->stream));
    }
    return mem ? HCX_SUCCESS : grow_item_max(c, qb, sp);
}

}  // namespace

extern "C" int hcx_last_timing(hcx_ctx *c, float *scan_ms, float *evalue_ms, float *total_ms) {
    if (!c) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_last_timing: null context");
    if (!c->timed_ok) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_last_timing: no timed search yet");
    float q = 0, e = 0, t = 0;
    HCX_HIP(hipEventElapsedTime(&q, c->ev0, c->ev1));
    HCX_HIP(hipEventElapsedTime(&e, c->ev4, c->ev5));
    HCX_HIP(hipEventElapsedTime(&t, c->ev2, c->ev3));
    if (scan_ms) *scan_ms = q;
    if (evalue_ms) *evalue_ms = e;
    if (total_ms) *total_ms = t;
    return HCX_SUCCESS;
}

extern "C" int hcx_last_scan_split(hcx_ctx *c, float *packed_ms, float *two_pass_ms) {
    if (!c) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_last_scan_split: null context");
    if (!c->timed_ok) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_last_scan_split: no timed search yet");
    float a = 0, b = 0;
    if (c->ran_packed) HCX_HIP(hipEventElapsedTime(&a, c->ev0, c->ev6));
    if (c->ran_two_pass) HCX_HIP(hipEventElapsedTime(&b, c->ev6, c->ev1));
    if (packed_ms) *packed_ms = a;
    if (two_pass_ms) *two_pass_ms = b;
    return HCX_SUCCESS;
}

extern "C" int hcx_surrogate_param_count(void) { return (int)mlp_param_count(); }

extern "C" int hcx_surrogate_init(hcx_ctx *c, uint64_t seed) {
    if (!c) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_surrogate_init: null context");
    HCX_HIP(hipSetDevice(c->device));
    // He-uniform weights (fan-in), zero biases; layout W1 b1 W2 b2 W3 b3
    const uint32_t P = mlp_param_count();
    std::vector<float> p(P, 0.0f);
    uint64_t st = seed ? seed : 0x5EED;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    };
    const uint32_t H = 128, F = 128, O = 16;
    size_t o = 0;
    const double a1 = std::sqrt(6.0 / F), a2 = std::sqrt(6.0 / H);
    for (uint32_t i = 0; i < H * F; i++) p[o++] = (float)(a1 * uni());
    o += H;
    for (uint32_t i = 0; i < H * H; i++) p[o++] = (float)(a2 * uni());
    o += H;
    for (uint32_t r = 0; r < O; r++)
        for (uint32_t k = 0; k < H; k++) p[o++] = r < 3 ? (float)(a2 * uni()) : 0.0f;
    const float norm[6] = {0, 0, 0, 1, 1, 1};
    return mlp_set(c, p.data(), norm);
}

extern "C" int hcx_surrogate_set(hcx_ctx *c, const float *params, uint32_t n, const float *norm6) {
    if (!c || !params || !norm6) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_surrogate_set: null argument");
    if (n != mlp_param_count()) return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_surrogate_set: wrong parameter count");
    HCX_HIP(hipSetDevice(c->device));
    return mlp_set(c, params, norm6);
}

extern "C" int hcx_surrogate_get(hcx_ctx *c, float *params, uint32_t n, float *norm6) {
    if (!c || !params || !norm6) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_surrogate_get: null argument");
    if (n != mlp_param_count()) return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_surrogate_get: wrong parameter count");
    HCX_HIP(hipSetDevice(c->device));
    return mlp_get(c, params, norm6);
}

extern "C" int hcx_surrogate_train(hcx_ctx *c, hcx_qbatch *qb, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                                   float *loss) {
    if (!c || !qb) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_surrogate_train: null argument");
    if (qb->ctx != c) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_surrogate_train: batch from another context");
    if (steps == 0 || !(lr > 0.0f)) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_surrogate_train: steps, lr");
    HCX_HIP(hipSetDevice(c->device));
    return mlp_train(c, qb, steps, batch, lr, seed, loss);
}

extern "C" int hcx_surrogate_gradient(hcx_ctx *c, const float *X, const float *T, uint32_t n, float *grad,
                                      float *loss) {
    if (!c || !X || !T || !grad || !loss) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_surrogate_gradient: null");
    HCX_HIP(hipSetDevice(c->device));
    return mlp_gradient(c, X, T, n, grad, loss);
}

extern "C" int hcx_surrogate_last_audit(const hcx_ctx *c, hcx_audit *out) {
    if (!c || !out) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_surrogate_last_audit: null argument");
    *out = c->last_audit;
    return HCX_SUCCESS;
}

extern "C" int hcx_selftest_mfma(hcx_ctx *c, const float *A, const float *B, float *D) {
    if (!c || !A || !B || !D) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_selftest_mfma: null argument");
    HCX_HIP(hipSetDevice(c->device));
    return selftest_mfma(c, A, B, D);
}

extern "C" int hcx_autotune(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb, int32_t kernel,
                            hcx_tuning *best) {
    if (!c || !qb || !best) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_autotune: null argument");
    if (kernel != HCX_KERNEL_BATCHED && kernel != HCX_KERNEL_PLAIN)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_autotune: tunes the batched or plain form");
    const uint32_t slots0 = c->slots, bpc0 = c->blocks_per_cu, grab0 = c->grab;
    struct Cand {
        uint32_t slots, bpc, grab;
    };
    std::vector<Cand> cands;
    const bool batched = kernel == HCX_KERNEL_BATCHED;
    if (c->scan_mode == HCX_SCAN_AUTO) {
        // the packed kernel's layout is fixed by the chunk size: the batched
        // form's grab size is what is left to choose
        if (batched)
            for (uint32_t g : {8u, 16u, 32u, 64u, 128u, 256u}) cands.push_back({slots0, bpc0, g});
        else
            cands.push_back({slots0, bpc0, grab0});
    } else {
        for (uint32_t sl : {512u, 1024u, 1536u, 2048u})
            for (uint32_t bp : {1u, 2u, 3u}) {
                if (batched)
                    for (uint32_t g : {32u, 64u, 128u}) cands.push_back({sl, bp, g});
                else
                    cands.push_back({sl, bp, grab0});
            }
    }
    hcx_search_opts o = {};
    o.kernel = kernel;
    o.evalue = HCX_EVAL_TAILFIT;
    hcx_tuning b = {};
    b.kernel_ms = -1.0f;
    auto restore = [&]() { c->slots = slots0, c->blocks_per_cu = bpc0, c->grab = grab0; };
    for (const Cand &cd : cands) {
        c->slots = cd.slots;
        c->blocks_per_cu = cd.bpc;
        c->grab = cd.grab;
        float k0 = 0, k1 = 0, t = 0;
        // two runs (the first also sizes the batch's item records); keep the better
        int rc = run_search(c, idx, nidx, qb, &o, true, &k0, &t);
        if (rc == HCX_ERR_INVLD_PARAM) continue;   // does not fit in LDS
        if (!rc) rc = run_search(c, idx, nidx, qb, &o, true, &k1, &t);
        if (rc) {
            restore();
            return rc;
        }
        const float k = std::min(k0, k1);
        if (b.kernel_ms < 0.0f || k < b.kernel_ms) {
            b.chunk = c->chunk;
            b.slots = cd.slots;
            b.blocks_per_cu = cd.bpc;
            b.hitlist = c->hitlist;
            b.grab = (int32_t)cd.grab;
            b.kernel_ms = k;
        }
    }
    if (b.kernel_ms < 0.0f) {
        restore();
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_autotune: no candidate fits");
    }
    c->slots = b.slots;
    c->blocks_per_cu = b.blocks_per_cu;
    c->grab = (uint32_t)b.grab;
    *best = b;
    return HCX_SUCCESS;
}

extern "C" int hcx_search_device(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb,
                                 const hcx_search_opts *opts, float *kernel_ms, float *total_ms) {
    if (!c) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_search_device: null context");
    return run_search(c, idx, nidx, qb, opts, true, kernel_ms, total_ms);
}

extern "C" int hcx_search_fetch(hcx_ctx *c, hcx_qbatch *qb, hcx_result *results, uint32_t *hist, hcx_stats *stats) {
    if (!c || !qb) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_search_fetch: null argument");
    HCX_HIP(hipSetDevice(c->device));
    HCX_HIP(hipStreamSynchronize(c->stream));
    const uint64_t n = (uint64_t)qb->nspec;
    if (!qb->res.p) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_search_fetch: no search has run");
    if (results && n) HCX_HIP(hipMemcpy(results, qb->res.p, sizeof(hcx_result) * n, hipMemcpyDeviceToHost));
    if (hist && n) HCX_HIP(hipMemcpy(hist, qb->hist.p, 4ull * c->hbins * n, hipMemcpyDeviceToHost));
    if (stats && n) HCX_HIP(hipMemcpy(stats, qb->stats.p, sizeof(hcx_stats) * n, hipMemcpyDeviceToHost));
    return HCX_SUCCESS;
}

extern "C" int hcx_search(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb,
                          const hcx_search_opts *opts, hcx_result *results, uint32_t *hist, hcx_stats *stats) {
    if (!c || !results) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_search: null argument");
    int rc = run_search(c, idx, nidx, qb, opts, false, nullptr, nullptr);
    if (rc) return rc;
    return hcx_search_fetch(c, qb, results, hist, stats);
}

extern "C" int hcx_search_noalloc(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb, int32_t kernel,
                                  hcx_result *d_results, uint32_t *d_hist, hcx_stats *d_stats) {
    if (!c || !d_results || !d_hist) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_search_noalloc: null argument");
    if (kernel != HCX_KERNEL_PLAIN && kernel != HCX_KERNEL_BATCHED && kernel != HCX_KERNEL_CURSOR)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_search_noalloc: unknown kernel");
    if (!c->scratch) return set_error_code(HCX_ERR_DEVICE, "hcx_search_noalloc: no scratch on this device");
    uint64_t nch_all = 0;
    int rc = search_shape(c, idx, nidx, qb, &nch_all);
    if (rc) return rc;
    SearchMem m;
    size_t need = 0;
    if (!search_mem_carve(c->scratch, kScratchBytes, (uint64_t)qb->nspec, nidx, nch_all, &m, &need))
        return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_search_noalloc: " + std::to_string(qb->nspec) + " spectra over " +
                                                      std::to_string(nidx) + " indices need " + std::to_string(need) +
                                                      " bytes of scratch; the limit is " +
                                                      std::to_string(kScratchBytes) + " (use hcx_search)");
    m.res = d_results;
    m.hist = d_hist;
    m.stats = d_stats;
    hcx_search_opts o = {};
    o.kernel = kernel;
    o.evalue = HCX_EVAL_TAILFIT;
    return run_search(c, idx, nidx, qb, &o, false, nullptr, nullptr, &m);
}

extern "C" int hcx_search_candidates(hcx_ctx *c, hcx_index *const *idx, int nidx, hcx_qbatch *qb, hcx_cand *out,
                                     uint64_t cap, uint64_t *n) {
    if (!c || !n || (cap && !out)) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_search_candidates: null argument");
    uint64_t nch_all = 0;
    int rc = search_shape(c, idx, nidx, qb, &nch_all);
    if (rc) return rc;
    HCX_HIP(hipSetDevice(c->device));
    SearchMem m;
    rc = search_mem_alloc(c, qb, nidx, nch_all, HCX_KERNEL_CURSOR, &m);
    if (rc) return rc;
    QueryArgs a;
    rc = prepare(c, idx, nidx, qb, nullptr, &a, m);
    if (rc) return rc;
    hcx::DeviceBuf dc, dn;
    if (dc.ensure(sizeof(hcx_cand) * cap + 16) || dn.ensure(16))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_search_candidates: alloc");
    HCX_HIP(hipMemsetAsync(dn.p, 0, 8, c->stream));
    a.cand = (hcx_cand *)dc.p;
    a.ncand = (unsigned long long *)dn.p;
    a.cand_cap = cap;
    rc = launch_candidates(c, a);
    if (rc) return rc;
    HCX_HIP(hipStreamSynchronize(c->stream));
    unsigned long long cnt = 0;
    HCX_HIP(hipMemcpy(&cnt, dn.p, 8, hipMemcpyDeviceToHost));
    *n = cnt;
    if (cap && cnt) HCX_HIP(hipMemcpy(out, dc.p, sizeof(hcx_cand) * std::min<uint64_t>(cap, cnt), hipMemcpyDeviceToHost));
    return HCX_SUCCESS;
}

extern "C" int hcx_tailfit_batch(hcx_ctx *c, const uint32_t *hist, uint32_t hbins, uint32_t n, const uint32_t *cpsms,
                                 const int32_t *maxhyp, int32_t *mu, int32_t *beta, int32_t *minhyp, int32_t *nexthyp,
                                 uint32_t *flags) {
    if (!c || (n && (!hist || !cpsms || !maxhyp || !mu || !beta || !minhyp || !nexthyp)))
        return set_error_code(HCX_ERR_INVLD_PTR, "hcx_tailfit_batch: null argument");
    if (hbins < 2 || hbins > 20002) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_tailfit_batch: bad hbins");
    for (uint32_t s = 0; s < n; s++)
        if (maxhyp[s] >= (int32_t)hbins)
            return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_tailfit_batch: maxhyp beyond histogram");
    if (n == 0) return HCX_SUCCESS;
    HCX_HIP(hipSetDevice(c->device));
    hcx::DeviceBuf dh, dc, dm, dout, dfl;
    if (dh.ensure(4ull * hbins * n) || dc.ensure(4ull * n) || dm.ensure(4ull * n) || dout.ensure(16ull * n) ||
        dfl.ensure(4ull * n))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_tailfit_batch: alloc");
    HCX_HIP(hipMemcpyAsync(dh.p, hist, 4ull * hbins * n, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemcpyAsync(dc.p, cpsms, 4ull * n, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemcpyAsync(dm.p, maxhyp, 4ull * n, hipMemcpyHostToDevice, c->stream));
    int rc = launch_tailfit_raw(c, (const uint32_t *)dh.p, hbins, n, (const uint32_t *)dc.p, (const int32_t *)dm.p,
                                (int32_t *)dout.p, (uint32_t *)dfl.p, c->force_host_fit);
    if (rc) return rc;
    HCX_HIP(hipStreamSynchronize(c->stream));
    std::vector<int32_t> o(4ull * n);
    HCX_HIP(hipMemcpy(o.data(), dout.p, 16ull * n, hipMemcpyDeviceToHost));
    std::vector<uint32_t> fl(n);
    HCX_HIP(hipMemcpy(fl.data(), dfl.p, 4ull * n, hipMemcpyDeviceToHost));
    std::vector<double> sx;
    for (uint32_t s = 0; s < n; s++) {
        if (fl[s] & HCX_RES_FIT_HOST) {
            // too close to a boundary for the device's log10: the host's
            // FP64 path decides
            if (sx.empty()) sx.resize(hbins);
            uint32_t f = 0;
            host_tailfit(hist + (size_t)s * hbins, hbins, cpsms[s], maxhyp[s], sx.data(), &o[4 * s], &f);
            fl[s] = f | HCX_RES_FIT_HOST;
        }
        mu[s] = o[4 * s];
        beta[s] = o[4 * s + 1];
        minhyp[s] = o[4 * s + 2];
        nexthyp[s] = o[4 * s + 3];
        if (flags) flags[s] = fl[s];
    }
    return HCX_SUCCESS;
}

extern "C" int hcx_part_export(hcx_ctx *c, hcx_qbatch *qb, int32_t qid0, uint32_t slot, hcx_part_res *packs,
                               uint16_t *slices) {
    if (!c || !qb || (qb->nspec > 0 && (!packs || (slot && !slices))))
        return set_error_code(HCX_ERR_INVLD_PTR, "hcx_part_export: null argument");
    if (slot == 0 || slot > 65536) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_part_export: slot in 1..65536");
    const int32_t n = qb->nspec;
    if (n == 0) return HCX_SUCCESS;
    if (!qb->res.p || !qb->hist.p) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_part_export: batch not searched");
    HCX_HIP(hipSetDevice(c->device));
    hcx::DeviceBuf dp, ds, dov;
    if (dp.ensure(sizeof(hcx_part_res) * (size_t)n) || ds.ensure(2ull * slot * (size_t)n) || dov.ensure(4))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_part_export: alloc");
    int rc = launch_fill_u32(c, (uint32_t *)dov.p, 1, 0);
    if (!rc)
        rc = launch_part_export(c, (const hcx_result *)qb->res.p, (const uint32_t *)qb->hist.p, c->hbins, n, qid0,
                                slot, (hcx_part_res *)dp.p, (uint16_t *)ds.p, (uint32_t *)dov.p);
    if (rc) return rc;
    uint32_t ov = 0;
    HCX_HIP(hipMemcpyAsync(&ov, dov.p, 4, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipMemcpyAsync(packs, dp.p, sizeof(hcx_part_res) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipMemcpyAsync(slices, ds.p, 2ull * slot * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    if (ov) return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_part_export: a histogram's occupied range exceeds slot");
    return HCX_SUCCESS;
}

extern "C" int hcx_combine(hcx_ctx *c, uint32_t nshards, uint32_t nspec, uint32_t slot,
                           const hcx_part_res *const *packs, const uint16_t *const *slices, hcx_fresult *out,
                           int32_t *keys) {
    if (!c || (nspec && (!packs || !slices || !out || !keys)))
        return set_error_code(HCX_ERR_INVLD_PTR, "hcx_combine: null argument");
    if (nshards == 0 || nshards > 4096 || slot == 0 || slot > 65536 || nspec > 0x7FFFFFFFu)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_combine: 1..4096 shards, slot in 1..65536");
    if (nspec == 0) return HCX_SUCCESS;
    for (uint32_t g = 0; g < nshards; g++) {
        if (!packs[g] || !slices[g]) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_combine: null shard");
        for (uint32_t s = 0; s < nspec; s++) {
            const hcx_part_res &r = packs[g][s];
            if (r.N >= 1 && (r.max2 < r.min || r.max2 >= c->hbins || (uint32_t)(r.max2 - r.min) >= slot))
                return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_combine: a slice outside the histogram or its slot");
        }
    }
    HCX_HIP(hipSetDevice(c->device));
    const size_t n = nspec;
    hcx::DeviceBuf dp, ds, dout, dkeys, dfix;
    if (dp.ensure(sizeof(hcx_part_res) * n * nshards) || ds.ensure(2ull * slot * n * nshards) ||
        dout.ensure(sizeof(hcx_fresult) * n) || dkeys.ensure(4 * n) || dfix.ensure(4 * (n + 1)))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_combine: alloc");
    for (uint32_t g = 0; g < nshards; g++) {
        HCX_HIP(hipMemcpyAsync((hcx_part_res *)dp.p + g * n, packs[g], sizeof(hcx_part_res) * n,
                               hipMemcpyHostToDevice, c->stream));
        HCX_HIP(hipMemcpyAsync((uint16_t *)ds.p + g * n * slot, slices[g], 2ull * slot * n, hipMemcpyHostToDevice,
                               c->stream));
    }
    int rc = launch_fill_u32(c, (uint32_t *)dfix.p, 1, 0);
    if (!rc)
        rc = launch_combine(c, (const hcx_part_res *)dp.p, (const uint16_t *)ds.p, nshards, (int32_t)nspec, slot,
                            (hcx_fresult *)dout.p, (int32_t *)dkeys.p, (uint32_t *)dfix.p);
    if (rc) return rc;
    uint32_t nfix = 0;
    HCX_HIP(hipMemcpyAsync(out, dout.p, sizeof(hcx_fresult) * n, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipMemcpyAsync(keys, dkeys.p, 4 * n, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipMemcpyAsync(&nfix, dfix.p, 4, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    if (nfix == 0) return HCX_SUCCESS;
    // fits too close to a boundary for the device's log10: the host merges
    // and fits those spectra again from the same inputs
    std::vector<uint32_t> ids(nfix);
    HCX_HIP(hipMemcpy(ids.data(), (uint32_t *)dfix.p + 1, 4ull * nfix, hipMemcpyDeviceToHost));
    std::vector<double> pd(c->hbins), sx(c->hbins);
    for (uint32_t s : ids) {
        std::fill(pd.begin(), pd.end(), 0.0);
        int32_t cpsms = 0, pN = 0, key = (int32_t)nshards;
        float maxhyp = -1.0f;
        for (uint32_t g = 0; g < nshards; g++) {
            const hcx_part_res &r = packs[g][s];
            if (r.min == 0 && r.max == 0.0f && r.N == 0 && r.max2 == 0) continue;
            cpsms += r.N;
            if (r.N >= 1) {
                pN += r.N;
                const uint16_t *buf = slices[g] + (size_t)s * slot;
                for (int jj = r.min; jj <= (int)r.max2; jj++) {
                    double v = buf[jj - r.min];
                    if (r.N > 65500) v = (v / 65500) * r.N;
                    pd[jj] = pd[jj] + v;
                }
                if (r.max > 0 && r.max > maxhyp) {
                    maxhyp = r.max;
                    key = (int32_t)g;
                }
            }
        }
        const int32_t hyp = (int32_t)((double)(maxhyp * 10.0f) + 0.5);
        const double e = hcx::host_tailfit_merged(pd.data(), c->hbins, pN, hyp, sx.data());
        // (int) as x86-64 converts: out of range gives INT_MIN
        auto i32_of = [](double v) { return (v > -2147483649.0 && v < 2147483648.0) ? (int32_t)v : INT32_MIN; };
        hcx_fresult f = {i32_of(c->params.expect_max * 1e6), -1, 0};
        int32_t k = (int32_t)nshards;
        if (e < c->params.expect_max) {
            f.eValue = i32_of(e * 1e6);
            f.specID = packs[key][s].qID;
            f.npsms = cpsms;
            k = key;
        }
        out[s] = f;
        keys[s] = k;
    }
    return HCX_SUCCESS;
}

// v_log_f32 of 1 .. n on the device (the packed kernel's score estimate).
extern "C" int hcx_selftest_log2(hcx_ctx *c, uint32_t n, float *out, double *bound) {
    if (!c || (n && !out)) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_selftest_log2: null argument");
    if (bound) *bound = kLog2Err;
    if (n == 0) return HCX_SUCCESS;
    HCX_HIP(hipSetDevice(c->device));
    hcx::DeviceBuf d;
    if (d.ensure(4ull * n)) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_selftest_log2: alloc");
    int rc = selftest_log2(c, n, (float *)d.p);
    if (rc) return rc;
    HCX_HIP(hipMemcpyAsync(out, d.p, 4ull * n, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    return HCX_SUCCESS;
}

// Host check of the divider the query kernels use for raw / speclen.
extern "C" int hcx_selftest_divider(uint32_t d, const uint32_t *nums, uint32_t count, uint32_t *out) {
    uint32_t m, s;
    make_divider(d, &m, &s);
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t n = nums[i];
        if (s == 0) {
            out[i] = n;
            continue;
        }
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
        out[i] = (t + ((n - t) >> 1)) >> (s - 1);
    }
    return HCX_SUCCESS;
}
