// hcx_internal.h -- device data layout and kernel launchers of libhcx.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hcx.h"

namespace hcx {

// Set the library's last error text and return HCX_ERR_DEVICE.
int set_error(const std::string &msg);
int set_error_code(int code, const std::string &msg);

#define HCX_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::hcx::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Bins are (uint)(m/z * scale) < NB = max_mass * scale.
#ifndef HCX_CHUNK_THREADS
#define HCX_CHUNK_THREADS 512
#endif
// threads per workgroup of the chunked scan kernels; a run-table group holds
// at most kProbesPerThread probes per thread
constexpr uint32_t kChunkThreads = HCX_CHUNK_THREADS;
constexpr uint32_t kProbesPerThread = 5;

// Per-length index as the query kernels read it.
struct LenDesc {
    const float *mass;      // n masses, ascending
    const uint32_t *iA;     // n * speclen ion ids sorted by bin
    const uint32_t *bA;     // NB + 1 bucket offsets
    const uint32_t *mdir;   // max_mass + 2 entries: first entry with mass >= d (Da)
    uint32_t n;
    uint32_t speclen;
    uint32_t half;
    uint32_t peplen;
    uint32_t div_m;         // magic multiplier for raw / speclen
    uint32_t div_s;         // shift
    uint32_t nch;           // peptide chunks (of ctx->chunk peptides)
    uint32_t cbase;         // first (length, chunk) work bucket of this length
    const uint32_t *bA2;    // start of chunk c inside bin b, at ba2_index(b, c, nch)
    const uint16_t *iAc;    // n * speclen: (peptide mod chunk) << 1 | isY, iA order
};

// Word of bA2 holding where chunk c (0 .. nch) starts inside bin b.  Bins
// are taken in groups of eight and a group's words are chunk-major, so the
// two words of a probe's run (chunks c and c + 1) are 32 bytes apart, one
// 128-byte line holds four consecutive chunks of eight bins (a peak's 2dF + 1
// bins touch one or two groups), and a spectrum's items on consecutive chunks
// probe lines an earlier item brought into L2.  Groups of 2, 4 and 16 bins
// and the plain chunk-major layout measured slower.
constexpr uint32_t kBa2Group = 8;
__host__ __device__ inline size_t ba2_index(uint32_t b, uint32_t c, uint32_t nch) {
    return ((size_t)(b / kBa2Group) * (nch + 1) + c) * kBa2Group + b % kBa2Group;
}
__host__ __device__ inline size_t ba2_words(uint32_t NB, uint32_t nch) {
    return (size_t)(NB + kBa2Group - 1) / kBa2Group * kBa2Group * (nch + 1);
}

struct DeviceBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);   // grow-only (hipMalloc); contents undefined after growth
    void release();
    ~DeviceBuf() { release(); }
};

}  // namespace hcx

struct hcx_ctx {
    hcx_params params;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux[2] = {nullptr, nullptr};   // default priority; pickpeaks size classes side by side
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    hipEvent_t ev4 = nullptr, ev5 = nullptr;   // around the e-value stage
    hipEvent_t ev6 = nullptr;                  // between the packed and the two-pass scan kernel
    bool ran_packed = false, ran_two_pass = false;   // scan kernels of the last timed search
    bool timed_ok = false;                     // the last search recorded all six events
    bool capturing = false;                    // launchers record no events (graph capture)
    double *d_lgf = nullptr;          // max_shdpeaks + 1
    std::vector<double> h_lgf;
    uint32_t NB = 0;                  // max_mass * scale
    uint32_t hbins = 0;               // 10 * max_hyperscore + 2
    int heap_cap = 0;
    int num_cu = 0;
    // scratch
    hcx::DeviceBuf lens;              // LenDesc[nidx]
    hcx::DeviceBuf tie;               // tie list: count + spectrum ids
    hcx::DeviceBuf work;              // batched kernel queue / counters
    hcx::DeviceBuf sur;               // surrogate parameters + optimizer state
    hcx::DeviceBuf sur_tmp;           // surrogate features / targets / outputs of a batch
    hcx::DeviceBuf mlp_buf;           // training activations
    float *mlp_loss = nullptr, *mlp_Y = nullptr;
    uint16_t *mlp_H1 = nullptr, *mlp_H2 = nullptr;
    uint32_t sur_step = 0;            // Adam steps taken
    hcx_audit last_audit = {};        // audit of the last surrogate search
    hcx::DeviceBuf scan_tmp;          // device scan scratch
    hcx::DeviceBuf bld[6];            // index build scratch: keys x2, values x2, digit counts, scan
    bool sur_ready = false;
    uint32_t chunk = 16384;           // peptides per query chunk (power of two)
    uint32_t slots = 1536;            // scored-peptide slots per chunk pass (measured best)
    uint32_t blocks_per_cu = 2;       // persistent grid of the batched kernel
    int32_t hitlist = -1;             // LDS hit-list entries: < 0 as many as fit
    uint32_t group = 0;               // probes per run table: 0 as many as fit, else 64..kProbesPerThread x threads
    int32_t scan_mode = HCX_SCAN_AUTO;   // packed scan where exact, two-pass elsewhere; or two-pass only
    uint32_t grab = 64;               // batched kernels: items per queue grab, at most
    uint32_t force_redo = 0;          // test hook (HCX_OPT_FORCE_REDO): re-run every n-th spectrum on the cursor kernel
    uint32_t force_host_fit = 0;      // test hook (HCX_OPT_FORCE_HOST_FIT): fit every n-th spectrum on the host
    int max_lds = 65536;              // device limit per workgroup
    void *scratch = nullptr;          // this device's copy of the allocation-free search's scratch array
    hcx::DeviceBuf fix_dev;           // host re-fit staging: kFixRows histogram rows, then kFixRows results
    void *fix_host = nullptr;         // pinned host copy of the same (+ 256 B for the count)
    // the captured search sequence and the launch shape it was captured for
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key = 0;
};

struct hcx_index {
    hcx_ctx *ctx = nullptr;
    uint32_t peplen = 0, speclen = 0, n = 0, nseqs = 0;
    uint64_t nions = 0;
    float build_ms = 0.0f;
    uint32_t chunk = 0, nch = 0;
    hcx::DeviceBuf entries, mass, iA, bA, mdir, seqs, bA2, iAc;
};

struct hcx_qbatch {
    hcx_ctx *ctx = nullptr;
    int32_t nspec = 0, file_num = 0;
    uint64_t npeaks = 0;
    uint32_t max_peaks = 0;
    uint32_t max_mult = 0;   // most peaks of one spectrum within +-dF of a single bin
    uint32_t max_int = 0;    // largest peak intensity of the batch
    hcx::DeviceBuf moz, inten, idx, prec, chg, rt;
    hcx::DeviceBuf res, hist, stats, feat, pred;
    hcx::DeviceBuf mult;                 // per spectrum: most peaks within 2 * dF + 1 bins
    hcx::DeviceBuf wsum;                 // per spectrum: largest intensity sum within 2 * dF + 1 bins (saturating)
    // chunked-search state: per-spectrum accumulators, windows, work-item
    // numbering, the two-pass kernel's list
    hcx::DeviceBuf acc, win, packed, counts, item_base, item_spec, item_max, fb, redo;
    uint32_t item_cap = 0;               // item_max entries (grown to the last search's item total)
};

namespace hcx {

// The batch facts of a batch in HBM (uploaded by hcx_queries_upload or made by
// pickpeaks), on the device (k_pick.hip): per-spectrum multiplicities and
// windowed intensity sums into b->mult and b->wsum, b->max_int, b->max_mult.
// Allocates b->mult and b->wsum.  Ends with the stream synchronised.
int batch_facts(hcx_ctx *c, hcx_qbatch *b, uint32_t maxp);

// Integer division n / d for all 32-bit n (round-up method with a 33-bit
// effective multiplier): q = (t + ((n - t) >> 1)) >> (s - 1), t = mulhi(m, n).
void make_divider(uint32_t d, uint32_t *m, uint32_t *s);

// ---- index build (k_index.hip) ----
struct BuildArgs {
    const char *seqs;
    const hcx_pep_entry *entries;
    uint32_t n, peplen, speclen, maxz, scale, NB;
    double min_mass, max_mass;
    uint32_t mod_mass[HCX_MAX_MOD_TYPES];
};
// Sizes the index build's scratch (kept by the context, grown geometrically)
// for nions ions; launch_index_build then allocates nothing.
int reserve_index_build(hcx_ctx *ctx, uint64_t nions, uint32_t NB);
int launch_index_build(hcx_ctx *ctx, const BuildArgs &a, uint32_t *d_iA, uint32_t *d_bA,
                       const float *d_mass, uint32_t *d_mdir, uint32_t max_mass);
// DSLIM_Construct into a new index from entries, masses and sequences in host
// (kind hipMemcpyHostToDevice) or device memory (hcx_api.hip).
int index_build_from(hcx_ctx *c, uint32_t peplen, const char *seqs, uint32_t nseqs, const hcx_pep_entry *entries,
                     const float *mass, uint32_t nentries, hipMemcpyKind kind, hcx_index **out);
int index_build_device(hcx_ctx *c, uint32_t peplen, const char *d_seqs, uint32_t nseqs, const hcx_pep_entry *d_entries,
                       const float *d_mass, uint32_t nentries, hcx_index **out);
// ---- peptide database stage (k_lbe.hip) ----
// Sorts the segments [seg[i].x, seg[i].y) of (K, V) by key into the order
// libstdc++'s std::sort leaves them; ends with the stream synchronised.
int sort_mass_segments(hcx_ctx *c, float *K, uint32_t *V, uint64_t total, const std::vector<uint2> &seg);

// ---- query (k_query.hip) ----
struct QueryArgs {
    const LenDesc *lens;
    int nlen;
    const uint32_t *moz, *inten, *qidx;
    const float *prec;
    int32_t nspec;
    uint32_t dF, NB, min_shp, hbins, max_shdpeaks, tile, nprobe_max, max_mass;
    double dM;
    const double *lgf;
    hcx_result *res;
    uint32_t *hist;
    hcx_stats *stats;
    uint32_t *tie;        // tie[0] = count, tie[1..] = spectrum ids
    hcx_cand *cand;       // candidate dump (optional)
    unsigned long long *ncand;
    uint64_t cand_cap;
    int heap_cap;
    uint32_t *queue;      // batched kernel: work counter
    const uint32_t *redo; // cursor kernel: [0] count, [1..] the spectra to (re)score; null: every spectrum
};
int launch_cursor(hcx_ctx *ctx, const QueryArgs &a);
int launch_replay_cursor(hcx_ctx *ctx, const QueryArgs &a);
// the cursor kernel over a.redo (persistent grid; the count is read on the device)
int launch_cursor_redo(hcx_ctx *ctx, const QueryArgs &a);
// ---- chunked search (k_chunk.hip) ----
struct SpecAcc {
    unsigned long long top_a, top_b, top_n;
    uint32_t cpsms, flags;
    unsigned long long n_ia, n_probe;
};
// Largest min_shp the chunked kernels' 8-bit saturating counters can test;
// a larger threshold runs on the cursor kernel.
constexpr uint32_t kMaxChunkShp = 64;
// Items per queue grab of the batched kernels, at most.
constexpr uint32_t kMaxGrab = 256;
// Work items of the batched kernels are (spectrum, length, chunk) triples,
// numbered spectrum by spectrum (item_base), then by length and chunk; they
// are never stored: a kernel maps an item number back to its triple through
// item_base, item_spec and the spectrum's windows.
struct ChunkArgs {
    QueryArgs q;
    uint32_t tc, slots, max_peaks, grid;
    uint32_t group;       // probes per run-table group
    uint32_t wdiv_m, wdiv_s;   // magic divider for 2 * dF + 1 (probes per peak)
    uint32_t hl_cap;      // LDS hit-list entries (0: none); needs fewer than 65535 peaks
    SpecAcc *acc;
    uint32_t *queue;
    uint32_t wide;        // 64-bit intensity sums in the slots
    uint32_t replay;      // layout of the in-order replay (slots keep their hyperscores)
    uint32_t grab;        // batched: items per queue grab, at most (<= kMaxGrab)
    const int2 *win;            // nspec * nlen precursor windows (lo, hi) in pepEntries rows; lo > hi: none
    const uint32_t *item_base;  // batched: nspec + 1, first item of each spectrum; [nspec] = item total
    const uint32_t *item_spec;  // batched: spectrum holding item 64 * i
    uint32_t *item_max;         // batched: each item's best hyperscore (float bits), or null
    uint32_t item_cap;          // entries of item_max: items past it are neither recorded nor skipped
    const uint8_t *packed;      // per spectrum: 1 = scored by the packed kernel
    const uint32_t *fb_list;    // spectra left to the two-pass kernel (null: every spectrum)
    const uint32_t *fb_count;   // their number (device)
    uint32_t *redo;             // [0] count, [1..]: spectra whose saturating counters neared overflow (k_finalize)
    uint32_t force_redo;        // test hook: every force_redo-th spectrum is listed as well (0: none)
};
// Inputs of the per-spectrum window kernel.
struct WinArgs {
    const uint32_t *mult;   // per spectrum: most peaks within 2 * dF + 1 bins
    const uint32_t *wsum;   // per spectrum: largest intensity sum within 2 * dF + 1 bins
    uint32_t half;          // ions per series of the longest length (speclen / 2)
    uint32_t pack;          // 1: spectra the packed kernel scores exactly go to it
    uint32_t batched;       // 1: count each listed spectrum's work items
    int2 *win;
    uint32_t *counts;       // nspec + 1 (batched)
    uint8_t *packed;
    uint32_t *fb_list, *fb_count;
    SpecAcc *acc;           // HCX_RES_TWO_PASS set for spectra the packed kernel leaves
};

size_t chunk_smem_bytes(const ChunkArgs &a);
int launch_chunk_offsets(hcx_ctx *ctx, const uint32_t *iA, const uint32_t *bA, uint32_t NB, uint32_t nch, uint32_t TC,
                         uint32_t speclen, uint32_t *bA2);
int launch_fill_u32(hcx_ctx *ctx, uint32_t *p, uint64_t n, uint32_t v);
int launch_chunk_local(hcx_ctx *ctx, const uint32_t *iA, uint64_t n, uint32_t speclen, uint32_t TC, uint16_t *iAc);
int launch_windows(hcx_ctx *ctx, const ChunkArgs &a, const WinArgs &w);
int launch_item_spec(hcx_ctx *ctx, const uint32_t *item_base, int32_t nspec, uint32_t *item_spec);
int launch_chunked(hcx_ctx *ctx, const ChunkArgs &a, int kernel);
int launch_finalize(hcx_ctx *ctx, const ChunkArgs &a);
// ---- packed single-pass search (k_packed.hip) ----
// threads per workgroup of the packed kernels and workgroups per CU (the
// accumulators of a chunk of TC peptides take 8 * TC bytes of the CU's LDS)
#ifndef HCX_PACKED_THREADS
#define HCX_PACKED_THREADS 1024
#endif
#ifndef HCX_PACKED_BLOCKS
#define HCX_PACKED_BLOCKS 1
#endif
constexpr uint32_t kPackedThreads = HCX_PACKED_THREADS;
constexpr uint32_t kPackedBlocks = HCX_PACKED_BLOCKS;
// Bound on |v_log_f32(n) - log2(n)| for the integers n = 1 .. 2^24 + 1 that the
// packed kernel's score estimate assumes (the device's measured error, from
// hcx_selftest_log2, must stay well below it).
constexpr double kLog2Err = 2.0e-5;
// The tail fit's log10 values: two correct-to-a-few-ulp log10s (the device's
// and a host libm's) differ by less than kFitYErr on the fit's range
// (|log10 v| < 10 for v in [2^-32, 1]: 4 ulp = 7.1e-15, with margin).
constexpr double kFitYErr = 3.0e-14;
// and a window threshold comparison is trusted when the value is further
// than kFitWinEps from the threshold
constexpr double kFitWinEps = 1.0e-12;
// the surrogate's located / plain-fitted counters: kSurSlots slots, each
// kSurStride words (a cache line) apart, so that one atomic add per spectrum
// does not queue on one address
constexpr uint32_t kSurSlots = 64, kSurStride = 32;
// spectra fitted again on the host per round (staging rows)
constexpr uint32_t kFixRows = 256;
int selftest_log2(hcx_ctx *ctx, uint32_t n, float *out);
// LDS the packed kernels need with a.group probes per run table, or 0 when
// a.tc peptides' accumulators do not fit beside the other arrays
size_t packed_smem_bytes(const ChunkArgs &a);
int launch_packed(hcx_ctx *ctx, const ChunkArgs &a, int kernel);
int launch_replay_chunked(hcx_ctx *ctx, ChunkArgs &a);
int scan_exclusive_u32(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tmp);
uint64_t scan_tmp_words(uint64_t n);

// fix: a list (count, then spectrum ids; room for n) that receives the
// spectra whose device fit sat too close to a truncation or window boundary
// to be trusted (HCX_RES_FIT_HOST); the host fits them again (fix_host_fits).
int launch_tailfit(hcx_ctx *ctx, hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                   uint32_t min_cpsm, double expect_max, uint32_t *fix, uint32_t force = 0);
// audit of surrogate e-values: the plain fit for every stride-th spectrum,
// compared with the surrogate's e-value; cnt[0] audited, cnt[1] disagreeing
// (relative error > tol), cnt[2..3] largest relative error (double bits).
int launch_tailfit_audit(hcx_ctx *ctx, hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                         uint32_t min_cpsm, double expect_max, uint32_t stride, double tol, uint32_t *cnt,
                         uint32_t *fix, uint32_t force = 0);
// plain-fit window ends (absolute bins, 2 per spectrum, -1: no regular fit)
int launch_fit_windows(hcx_ctx *ctx, const hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                       uint32_t min_cpsm, int32_t *win);
// surrogate e-values: predicted window ends located exactly, else the plain
// fit; cnt: kSurSlots slots of (located, plain) kSurStride words apart
int launch_sur_fit(hcx_ctx *ctx, const float *Y, uint32_t ystride, const float *norm, hcx_result *res,
                   const uint32_t *hist, uint32_t hbins, int32_t n, uint32_t min_cpsm, double expect_max,
                   uint32_t *fix, uint32_t force, uint32_t *cnt);
// shard partial results (one wave per spectrum; *overflow set when a
// histogram's occupied range exceeds slot) and their merge (fix: the host
// re-fit list, count then ids)
int launch_part_export(hcx_ctx *ctx, const hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                       int32_t qid0, uint32_t slot, hcx_part_res *packs, uint16_t *slices, uint32_t *overflow);
int launch_combine(hcx_ctx *ctx, const hcx_part_res *packs, const uint16_t *slices, uint32_t G, int32_t n,
                   uint32_t slot, hcx_fresult *out, int32_t *keys, uint32_t *fix);
// ---- host tail fit (fit_host.hip) ----
// expeRT::ModelTailFit on the host in FP64 with the host libm's log10, for
// the spectra the device lists as fragile.  sx: hbins doubles of workspace;
// out4: mu, beta, stt1, end1.
void host_tailfit(const uint32_t *H, uint32_t hbins, uint32_t cpsms, int32_t hyp, double *sx, int32_t out4[4],
                  uint32_t *flags);
// bytes of the pinned host staging: count, kFixRows rows and results, and
// one fit's workspace
inline size_t fix_host_bytes(uint32_t hbins) {
    return 256 + (size_t)kFixRows * hbins * 4 + (size_t)kFixRows * sizeof(hcx_result) + (size_t)hbins * 8;
}
// expeRT::ModelTailFit(double&, int) on a merged histogram, on the host
// (sx: hbins doubles of workspace); returns the e-value.
double host_tailfit_merged(const double *pd, uint32_t hbins, int32_t pN, int32_t hyp, double *sx);
// Fits the spectra of the device list `fix` again on the host and writes
// their results back (kFixRows at a time through the context's staging
// buffers); ends with the stream synchronised.
int fix_host_fits(hcx_ctx *c, const uint32_t *fix, hcx_result *res, const uint32_t *hist, uint32_t hbins,
                  double expect_max);
int launch_fix_gather(hcx_ctx *c, const uint32_t *ids, uint32_t k, const uint32_t *hist, uint32_t hbins,
                      const hcx_result *res, uint32_t *rows, hcx_result *rstage);
int launch_fix_scatter(hcx_ctx *c, const uint32_t *ids, uint32_t k, const hcx_result *rstage, hcx_result *res);
// ---- surrogate (k_mlp.hip) ----
uint32_t mlp_param_count();
int mlp_alloc(hcx_ctx *c);
int mlp_set(hcx_ctx *c, const float *params, const float *norm6);
int mlp_get(hcx_ctx *c, float *params, float *norm6);
// features, network, then e-values by k_sur_fit (fix: the host re-fit list;
// cnt: the located / plain-fit counter slots)
int mlp_infer(hcx_ctx *c, hcx_qbatch *qb, uint32_t *fix, uint32_t *cnt);
int mlp_gradient(hcx_ctx *c, const float *X, const float *T, uint32_t n, float *grad, float *loss);
int mlp_train(hcx_ctx *c, hcx_qbatch *qb, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float *loss_out);
int selftest_mfma(hcx_ctx *c, const float *A, const float *B, float *D);

int launch_tailfit_raw(hcx_ctx *ctx, const uint32_t *hist, uint32_t hbins, uint32_t n, const uint32_t *cpsms,
                       const int32_t *maxhyp, int32_t *out4, uint32_t *flags, uint32_t force = 0);

}  // namespace hcx
