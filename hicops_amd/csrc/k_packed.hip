// k_packed.hip -- single-pass shared-peak scan with packed per-peptide
// accumulators: the scan for every spectrum it can score exactly.
//
// Work item = (spectrum, length, chunk of TC peptides), as in k_chunk.hip, on
// the same (bin, chunk)-bucketed index (bA2 runs, 16-bit chunk-local iAc).
// Each peptide of the chunk owns two 32-bit LDS words, one per ion series:
// word index = the iAc entry itself (local << 1 | isY), word = count << 24 |
// intensity sum.  For one item:
//   R  every (peak, bin) probe reads its run [bA2[c], bA2[c+1]); the
//      non-empty runs are compacted into one flat sequence whose run records
//      carry (entry index minus flat offset, (1 << 24) | the probe's peak
//      intensity), with a bitmap of run ends (as k_chunk.hip);
//   P  one pass over the flat entries: lane j of a 64-entry tile reads entry
//      F + j (coalesced 16-bit loads) and, if the peptide is inside the
//      precursor window (dslim_query.cpp:637-641), adds the run's value to
//      the entry's word -- one non-returning LDS atomic per entry, which
//      counts the shared peak and sums its intensity at once
//      (dslim_query.cpp:644-676);
//   S  the window's words are read 128 bits at a time and cleared; every
//      peptide with bc + yc >= min_shp is scored (dslim_query.cpp:679-722):
//      hs = lgf[bc] + lgf[yc] + log10(1+ibc) + log10(1+iyc) - 6 in FP64,
//      rounded to float, histogram bin (int)(hs*10f + 0.5), the spectrum's
//      best.
// Per-spectrum results merge across workgroups as in k_chunk.hip (cpsms and
// histogram atomics, hs-keyed 64-bit maxima); tied maxima go to the in-order
// heap replay there.
//
// Exactness: the peaks that hit one ion bin lie within dF of it, in one
// window of 2 * dF + 1 bins, so they number at most M (the spectrum's most
// peaks within 2 * dF + 1 bins) and their intensities sum to at most W (its
// largest intensity sum within 2 * dF + 1 bins); a series of speclen / 2 ions
// takes at most speclen / 2 * M hits summing to at most speclen / 2 * W.  A
// spectrum is marked packed (k_windows) only if those stay below 256 and
// 2^24 for the longest length: no field can carry into the next.  The other
// spectra run on the two-pass kernel (k_chunk.hip), whose counts and sums
// are separate.
//
// A workgroup is 1024 threads (16 wavefronts, four per SIMD) and takes most
// of the CU's 160 KB of LDS: 16384 peptides x 8 bytes of accumulators plus the
// run table.
#include "k_scan_common.h"

namespace hcx {
namespace {

constexpr int kPT = (int)kPackedThreads;
// run-table probes per thread: a group holds at most kPPp * kPT probes
constexpr uint32_t kPPp = 2;
constexpr uint32_t kCnt = 1u << 24;   // count unit of a packed word
constexpr double kLog10of2 = 0.30102999566398119521;
// tiles per wavefront step and accumulator reads in flight per thread in the
// scoring scan: 4 and 2 measured fastest (tiles 2, 3, 6, 8, 12, 16 and reads
// 1, 4, 8 slower on the 10 M-peptide open search)
#ifndef HCX_PACKED_TILE_U
#define HCX_PACKED_TILE_U 4
#endif
constexpr uint32_t kUp = HCX_PACKED_TILE_U;
constexpr uint32_t kSr = 2;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef AS_LDS u32x2 lu32x2;

// Byte offsets of the dynamic LDS arrays (host and device agree on them).
struct PLayout {
    uint32_t acc;       // 2 * tc words at offset 0: b and y word of each peptide
    uint32_t runs;      // group + 1 uint2: (entry index - flat offset, (1 << 24) | intensity)
    uint32_t cofs;      // group + 1: flat offset of each non-empty run (exclusive prefix)
    uint32_t pk_moz, pk_int;   // max_peaks
    uint32_t seg;       // kSegWords u64: run ends of the current segment
    uint32_t wsum;      // 2 * (kPT / 64) + 2
    uint32_t hist;      // min(hbins, kHistLds)
    uint32_t lgf;       // max_shdpeaks + 1 doubles
    uint32_t wlo, whi;  // nlen
    uint32_t lfirst;    // nlen + 1
    uint32_t ctl;       // 8 words: [0] best exact score of the spectrum, [2..5] spectrum merge, [6] grabbed item
    uint32_t imax;      // kMaxGrab
    uint32_t red64;     // 2 u64
    uint32_t total;
};

__host__ __device__ inline uint32_t pal16(uint32_t x) { return (x + 15) & ~15u; }

__host__ __device__ inline PLayout make_playout(const ChunkArgs &a) {
    PLayout l;
    uint32_t off = 0;
    auto take = [&](uint32_t bytes) {
        uint32_t o = off;
        off = pal16(off + bytes);
        return o;
    };
    l.acc = take(8u * a.tc);
    l.runs = take(8u * (a.group + 1));
    l.cofs = take(4u * (a.group + 1));
    l.pk_moz = take(4u * a.max_peaks);
    l.pk_int = take(4u * a.max_peaks);
    l.seg = take(8u * kSegWords);
    l.wsum = take(4u * (2 * (kPT / 64) + 2));
    l.hist = take(4u * min(a.q.hbins, kHistLds));
    l.lgf = take(8u * (a.q.max_shdpeaks + 1));
    l.wlo = take(4u * a.q.nlen);
    l.whi = take(4u * a.q.nlen);
    l.lfirst = take(4u * (a.q.nlen + 1));
    l.ctl = take(32u);
    l.imax = take(4u * kMaxGrab);
    l.red64 = take(16u);
    l.total = off;
    return l;
}

struct PSmem {
    lu32 *acc, *cofs, *pk_moz, *pk_int, *wsum, *hist, *lfirst, *ctl, *imax;
    uint32_t cand_cap;   // candidate list entries (12 bytes) the run table's space holds
    lu32 *hbest;         // best exact score of the current spectrum so far (float bits)
    lu32x2 *runs;
    lu32 *seg32;
    lu64 *seg64, *red64;
    lf64 *lgf;
    li32 *wlo, *whi;
};

__device__ __forceinline__ PSmem pbind(unsigned char *raw, const PLayout &l) {
    AS_LDS unsigned char *b = (AS_LDS unsigned char *)raw;
    PSmem s;
    s.acc = (lu32 *)(b + l.acc);
    s.runs = (lu32x2 *)(b + l.runs);
    s.cofs = (lu32 *)(b + l.cofs);
    s.pk_moz = (lu32 *)(b + l.pk_moz);
    s.pk_int = (lu32 *)(b + l.pk_int);
    s.seg32 = (lu32 *)(b + l.seg);
    s.seg64 = (lu64 *)(b + l.seg);
    s.wsum = (lu32 *)(b + l.wsum);
    s.hist = (lu32 *)(b + l.hist);
    s.lgf = (lf64 *)(b + l.lgf);
    s.wlo = (li32 *)(b + l.wlo);
    s.whi = (li32 *)(b + l.whi);
    s.lfirst = (lu32 *)(b + l.lfirst);
    s.ctl = (lu32 *)(b + l.ctl);
    s.imax = (lu32 *)(b + l.imax);
    s.red64 = (lu64 *)(b + l.red64);
    s.cand_cap = (l.pk_moz - l.runs) / 12;
    s.hbest = s.ctl;
    return s;
}

// Run table of the probes [g0, g1) of one item (at most kPPp * kPT probes):
// each (peak, bin) probe with dF < q < NB-1-dF (dslim_query.cpp:625) reads
// its run [bA2[c], bA2[c+1]) in chunk c; the non-empty runs are compacted in
// probe order.  Returns (E, nc): flat entries and runs, uniform.
__device__ __forceinline__ uint2 setup_runs_p(const ChunkArgs &a, const PSmem &sm, const LenDesc &L, uint32_t c,
                                              uint32_t g0, uint32_t g1, bool count_probes, Acc &acc) {
    const uint32_t W = 2 * a.q.dF + 1;
    const uint32_t n = g1 - g0;
    uint32_t rs[kPPp], len[kPPp], pk[kPPp];
    uint32_t nr = 0, ne = 0;
#pragma unroll
    for (uint32_t h = 0; h < kPPp; h++) {
        rs[h] = len[h] = pk[h] = 0;
        const uint32_t lp = h * kPT + threadIdx.x;
        if (lp < n) {
            const uint32_t p = g0 + lp;
            const uint32_t k = udiv(p, a.wdiv_m, a.wdiv_s), j = p - k * W;
            pk[h] = k;
            const uint32_t q = sm.pk_moz[k];
            if (q > a.q.dF && q < a.q.NB - 1 - a.q.dF) {
                const size_t o = ba2_index(q - a.q.dF + j, c, L.nch);
                rs[h] = ldg(L.bA2, o);
                len[h] = ldg(L.bA2, o + kBa2Group) - rs[h];
                if (count_probes) acc.nprobe++;
            }
        }
        nr += len[h] ? 1u : 0u;
        ne += len[h];
    }
    // the bitmap is cleared before the scan's barrier and filled after it
    for (uint32_t w = threadIdx.x; w < 2 * kSegWords; w += kPT) sm.seg32[w] = 0;
    uint2 tot;
    const uint2 pre = block_prefix2<kPT>(sm.wsum, nr, ne, &tot);
    uint32_t ci = pre.x, fo = pre.y;
#pragma unroll
    for (uint32_t h = 0; h < kPPp; h++) {
        if (len[h]) {
            sm.runs[ci] = (u32x2){rs[h] - fo, kCnt | sm.pk_int[pk[h]]};
            sm.cofs[ci] = fo;
            fo += len[h];
            if (fo - 1 < kSegEntries) lds_or(&sm.seg32[(fo - 1) >> 5], 1u << ((fo - 1) & 31u));
            ci++;
        }
    }
    if (threadIdx.x == 0) {
        sm.cofs[tot.x] = tot.y;
        sm.runs[tot.x] = (u32x2){0u, 0u};
    }
    __syncthreads();
    return make_uint2(ufl(tot.y), ufl(tot.x));
}

// Score one peptide from its two packed words (dslim_query.cpp:679-722):
//   hs = (float)(((lgf[bc] + lgf[yc]) + log10(1 + ibc)) + log10(1 + iyc) - 6)
// in FP64 (the exact path).  Most candidates need only three facts about hs:
// its sign, its histogram bin and that it is below the spectrum's best; an
// estimate with a proven error bound settles them without the two FP64
// log10: log2 of the exact integer arguments (at most 2^24, exact in float)
// by the hardware's v_log_f32, whose largest error over 1 .. 2^24 + 1 is
// below kLog2Err (hcx_selftest_log2 measures it on the device).  The exact
// path runs whenever the estimate's interval [ha - E, ha + E] leaves a fact
// open: the sign, a bin boundary within reach, or a score that could reach the
// best exact score of the spectrum seen so far (*hbest, float bits shared by
// the workgroup: a stale value is lower, so only safer).  The spectrum's
// maximum and every peptide tied with it are therefore always exact.
__device__ __forceinline__ void score_peptide(const ChunkArgs &a, const PSmem &sm, uint32_t spec, uint32_t ixx,
                                              uint32_t pid, uint32_t wb, uint32_t wy, Acc &acc, float &imax) {
    const uint32_t bc = wb >> 24, yc = wy >> 24;
    if ((uint16_t)(bc + yc) < a.q.min_shp) return;
    const uint32_t ib = wb & (kCnt - 1), iy = wy & (kCnt - 1);
    uint32_t bi = bc, yi = yc;
    if (bi > a.q.max_shdpeaks) { bi = a.q.max_shdpeaks; acc.flags |= HCX_RES_LGF_CLAMP; }
    if (yi > a.q.max_shdpeaks) { yi = a.q.max_shdpeaks; acc.flags |= HCX_RES_LGF_CLAMP; }
    const double lg = sm.lgf[bi] + sm.lgf[yi];
    // the estimate and its error bound (the FP64 roundings of both paths
    // are below 1e-12 at these magnitudes)
    const double l2 = (double)__builtin_amdgcn_logf((float)(1u + ib)) + (double)__builtin_amdgcn_logf((float)(1u + iy));
    const double ha = (lg + l2 * kLog10of2) - 6.0;
    const double E = 2.0 * kLog2Err * kLog10of2 + 1e-12;
    const double hi = ha + E;
    if (!(hi > 0.0)) return;   // hs <= 0: not a candidate
    const float H = __uint_as_float(*sm.hbest);
    // fl(D) <= D + 2^-24 D: below H for sure only if hi (1 + 2^-23) < H
    bool exact = !(ha - E > 0.0) || hi * (1.0 + 0x1p-23) >= (double)H;
    int bin = 0;
    if (!exact) {
        // bin = floor(fl(fl(hs) * 10) + 0.5); the two float roundings move
        // 10 hs by at most 2^-23 of it
        const double z = 10.0 * ha, M = 10.0 * E + 0x1p-22 * z;
        const double b0 = floor(z - M + 0.5), b1 = floor(z + M + 0.5);
        if (b0 != b1) exact = true;
        else bin = (int)b0;
    }
    float hs;
    if (exact) {
        double t = lg;
        t = t + log10((double)(1u + ib));
        t = t + log10((double)(1u + iy));
        t = t - 6.0;
        hs = (float)t;
        if (!(hs > 0.0f)) return;
        bin = (int)((double)(hs * 10.0f) + 0.5);
        if (hs > H) lds_max(sm.hbest, __float_as_uint(hs));
        if (hs > acc.best_hs) {
            acc.best_hs = hs;
            acc.best_cnt = 1;
            acc.best_id = (ixx << 25) | pid;
            acc.best_by = (bc << 16) | yc;
        } else if (hs == acc.best_hs) {
            acc.best_cnt++;
        }
    } else {
        // an upper bound of hs, for the item's best (the tie replay only
        // skips items whose best is below the heap's root)
        hs = (float)(hi * (1.0 + 1e-6));
    }
    acc.cpsms++;
    imax = fmaxf(imax, hs);
    if (bin >= (int)a.q.hbins) { bin = (int)a.q.hbins - 1; acc.flags |= HCX_RES_HIST_CLAMP; }
    if ((uint32_t)bin < kHistLds) lds_add(&sm.hist[bin], 1u);
    else glb_add(a.q.hist + (size_t)spec * a.q.hbins + bin, 1u);
}

// One (spectrum, length, chunk) item; the spectrum's peaks are in LDS.
// imax_out (batched form): the item's best score, LDS max of float bits.
__device__ __forceinline__ void process_item_p(const ChunkArgs &a, const PSmem &sm, uint32_t spec, lu32 *imax_out,
                                               uint32_t P, uint32_t ixx, uint32_t c, int32_t lo, int32_t hi,
                                               Acc &acc) {
    const LenDesc L = load_len_u(a, ixx);
    const uint32_t TC = a.tc;
    const uint32_t base = c * TC;
    const uint32_t ilo = max((uint32_t)lo, base);
    const uint32_t ihi = min((uint32_t)hi, base + TC - 1);
    const uint32_t wlo = ilo - base, wspan = ihi - ilo;   // window in chunk-local ids
    const bool first = (uint32_t)lo / TC == c;            // probes are counted once per length
    const uint32_t NP = P * (2 * a.q.dF + 1);
    const uint32_t GP = a.group;
    const uint32_t ng = (NP + GP - 1) / GP;
    // iAc through a buffer resource: 32-bit offsets, and reads past the
    // array return 0 instead of faulting
    const __amdgpu_buffer_rsrc_t rs_iac =
        __builtin_amdgcn_make_buffer_rsrc((void *)L.iAc, 0, (int)(L.n * L.speclen * 2u), kBufDword3);
    const uint32_t lane = threadIdx.x & 63u;
    const bool full = wlo == 0 && wspan == TC - 1;
    uint32_t nia_w = 0;
    auto pass = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        for (uint32_t g = 0; g < ng; g++) {
            const uint32_t g0 = g * GP, g1 = min(NP, g0 + GP);
            const uint2 en = setup_runs_p(a, sm, L, c, g0, g1, first, acc);
            const uint32_t E = en.x, nc = en.y;
            if (FULL && threadIdx.x == 0) nia_w += E;   // per wavefront otherwise
            for (uint32_t s0 = 0; s0 < E; s0 += kSegEntries) {
                if (s0) build_seg<kPT>(sm.seg32, sm.cofs, nc, s0);
                uint32_t cv[kUp];
                walk_tiles<kPT, kUp>(
                    sm.seg64, sm.cofs, nc, s0, min(E, s0 + kSegEntries),
                    [&](const Tile (&t)[kUp], const uint32_t (&x)[kUp], const bool (&)[kUp],
                        uint32_t (&v)[kUp], uint32_t) {
                        uint32_t gi[kUp];
#pragma unroll
                        for (uint32_t u = 0; u < kUp; u++) {
                            const u32x2 r = (sm.runs + t[u].C)[tile_below(t[u])];
                            gi[u] = r.x + x[u];
                            cv[u] = r.y;
                        }
                        // lanes outside the share read some entry or, past
                        // the array, 0 (buffer range check); they add nothing
#pragma unroll
                        for (uint32_t u = 0; u < kUp; u++) v[u] = iac_load(rs_iac, gi[u]);
                    },
                    [&](const Tile (&)[kUp], const uint32_t (&v)[kUp], const uint32_t (&)[kUp],
                        const bool (&ok)[kUp], uint32_t) {
#pragma unroll
                        for (uint32_t u = 0; u < kUp; u++) {
                            const bool win = FULL ? ok[u] : ok[u] && (v[u] >> 1) - wlo <= wspan;
                            if (!FULL) nia_w += (uint32_t)__popcll(__ballot(win));
                            if (win) lds_add(&sm.acc[v[u]], cv[u]);
                        }
                    });
            }
            // the next group's set-up rewrites the run table and the bitmap
            __syncthreads();
        }
    };
    if (full) pass(std::true_type{});
    else pass(std::false_type{});
    if (lane == 0) acc.nia += nia_w;

    // S: the window's words, two peptides per 128-bit exchange with zeros
    // (read and cleared for the next item, whose pass starts after a barrier).  Peptides
    // with bc + yc >= min_shp are few and scattered, so each wavefront queues
    // its own in a private part of the run table (free by now; no atomics)
    // and scores them 64 at a time, instead of taking the scoring branch for
    // one or two lanes of every read.  The last probe group's barrier has
    // every wavefront's atomics done; an item without probes had no pass, and
    // needs a barrier of its own before its words are read (a wavefront may
    // still be clearing them in the previous item's scan).
    if (ng == 0) __syncthreads();
    float imax = 0.0f;
    {
        const uint32_t w = threadIdx.x >> 6;
        const uint32_t qcap = sm.cand_cap / (kPT / 64);   // queue entries per wavefront
        lu32 *c_pid = (lu32 *)sm.runs + w * 3 * qcap, *c_wb = c_pid + qcap, *c_wy = c_wb + qcap;
        uint32_t qn = 0;   // queued, uniform
        auto drain = [&]() {
            __builtin_amdgcn_wave_barrier();
            for (uint32_t i = lane; i < qn; i += 64) score_peptide(a, sm, spec, ixx, base + c_pid[i], c_wb[i], c_wy[i], acc, imax);
            __builtin_amdgcn_wave_barrier();
            qn = 0;
        };
        const uint32_t q0 = wlo >> 1, q1 = (wlo + wspan) >> 1;
        // a peptide without hits scores -6 and is never a candidate, so the
        // count test may use max(min_shp, 1); counts are at most 255 each, so
        // bc + yc is the reference's unsigned short sum
        const uint32_t shp = max(a.q.min_shp, 1u);
        // kSr reads in flight per thread, then their candidates; every read
        // word is cleared (no test: the writes are cheaper than the branch)
        for (uint32_t qb = q0; qb <= q1; qb += kSr * kPT) {
            u32x4 w4[kSr];
#pragma unroll
            for (uint32_t r = 0; r < kSr; r++) {
                const uint32_t qi = qb + r * kPT + threadIdx.x;
                w4[r] = (u32x4){0u, 0u, 0u, 0u};
                if (qi <= q1) {
                    // read and clear in one instruction (ds_wrxchg_rtn_b64)
                    lu64 *pw = (lu64 *)sm.acc + 2 * qi;
                    const unsigned long long lo2 =
                        __hip_atomic_exchange(pw, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const unsigned long long hi2 =
                        __hip_atomic_exchange(pw + 1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    w4[r] = (u32x4){(uint32_t)lo2, (uint32_t)(lo2 >> 32), (uint32_t)hi2, (uint32_t)(hi2 >> 32)};
                }
            }
#pragma unroll
            for (uint32_t r = 0; r < kSr; r++) {
                const uint32_t qi = qb + r * kPT + threadIdx.x;
                const u32x4 wv = w4[r];
                const bool k0 = (wv.x >> 24) + (wv.y >> 24) >= shp;
                const bool k1 = (wv.z >> 24) + (wv.w >> 24) >= shp;
                const unsigned long long m0 = __ballot(k0), m1 = __ballot(k1);
                const uint32_t tot = (uint32_t)(__popcll(m0) + __popcll(m1));
                if (tot == 0) continue;   // uniform per wavefront
                if (qn + tot > qcap) drain();
                // entries past the queue's end (a queue shorter than this
                // read's candidates) are scored in place
                uint32_t slot = qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(m0 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m0, 0u)) +
                                __builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
                if (k0) {
                    if (slot < qcap) {
                        c_pid[slot] = 2 * qi;
                        c_wb[slot] = wv.x;
                        c_wy[slot] = wv.y;
                    } else {
                        score_peptide(a, sm, spec, ixx, base + 2 * qi, wv.x, wv.y, acc, imax);
                    }
                    slot++;
                }
                if (k1) {
                    if (slot < qcap) {
                        c_pid[slot] = 2 * qi + 1;
                        c_wb[slot] = wv.z;
                        c_wy[slot] = wv.w;
                    } else {
                        score_peptide(a, sm, spec, ixx, base + 2 * qi + 1, wv.z, wv.w, acc, imax);
                    }
                }
                qn = min(qn + tot, qcap);
            }
        }
        if (qn) drain();
    }
    // the item's best score (one LDS atomic per wavefront: same-address
    // atomics serialise)
    if (imax_out) {
        const uint32_t wm = wave_max_last(__float_as_uint(imax));
        if (lane == 63u && wm) lds_max(imax_out, wm);
    }
}

__device__ __forceinline__ void init_psmem(const ChunkArgs &a, const PSmem &sm) {
    for (uint32_t i = threadIdx.x; i <= a.q.max_shdpeaks; i += kPT) sm.lgf[i] = a.q.lgf[i];
    for (uint32_t i = threadIdx.x; i < a.tc / 2; i += kPT) ((lu32x4 *)sm.acc)[i] = (u32x4){0u, 0u, 0u, 0u};
    for (uint32_t i = threadIdx.x; i < min(a.q.hbins, kHistLds); i += kPT) sm.hist[i] = 0;
    if (threadIdx.x < 8) sm.ctl[threadIdx.x] = 0;
}

// Plain form: one workgroup per spectrum (those marked packed), its items in
// order.
__global__ __launch_bounds__(kPT, 4 * kPackedBlocks * kPT / 1024) void k_spec_packed(ChunkArgs a) {
    const uint32_t s = blockIdx.x;
    if (!a.packed[s]) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const PSmem sm = pbind(smem_raw, make_playout(a));
    init_psmem(a, sm);
    const uint32_t P = load_peaks<kPT>(a, sm.pk_moz, sm.pk_int, s);
    for (int l = threadIdx.x; l < a.q.nlen; l += kPT) {
        const int2 w = a.win[(size_t)s * a.q.nlen + l];
        sm.wlo[l] = w.x;
        sm.whi[l] = w.y;
    }
    __syncthreads();
    Acc acc;
    acc_reset(acc);
    for (uint32_t ixx = 0; ixx < (uint32_t)a.q.nlen; ixx++) {
        const int32_t lo = ufl((int32_t)sm.wlo[ixx]), hi = ufl((int32_t)sm.whi[ixx]);
        if (lo > hi) continue;
        for (uint32_t c = (uint32_t)lo / a.tc; c <= (uint32_t)hi / a.tc; c++)
            process_item_p(a, sm, s, nullptr, P, ixx, c, lo, hi, acc);
    }
    flush_spectrum<kPT>(a, sm.hist, s, acc, sm.ctl + 2, sm.red64);
}

// Batched form: persistent workgroups pull (spectrum, length, chunk) items of
// the packed spectra (run_items).
__global__ __launch_bounds__(kPT, 4 * kPackedBlocks * kPT / 1024) void k_items_packed(ChunkArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const PSmem sm = pbind(smem_raw, make_playout(a));
    init_psmem(a, sm);
    Acc acc;
    acc_reset(acc);
    uint32_t P = 0;
    run_items<kPT>(
        a, sm.ctl + 6, sm.imax, sm.wlo, sm.lfirst,
        [&](uint32_t s) {
            P = load_peaks<kPT>(a, sm.pk_moz, sm.pk_int, s);
            load_windows<kPT>(a, s, sm.wlo, sm.whi, sm.lfirst);
            if (threadIdx.x == 0) *sm.hbest = 0u;
        },
        [&](uint32_t s, uint32_t l, uint32_t c, lu32 *imax_out) {
            process_item_p(a, sm, s, imax_out, P, l, c, ufl((int32_t)sm.wlo[l]), ufl((int32_t)sm.whi[l]), acc);
        },
        [&](uint32_t s) { flush_spectrum<kPT>(a, sm.hist, s, acc, sm.ctl + 2, sm.red64); });
}

// log2 of the integers 1 .. n by v_log_f32, as the score estimate computes it
__global__ void k_selftest_log2(uint32_t n, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = __builtin_amdgcn_logf((float)(i + 1u));
}

}  // namespace

int selftest_log2(hcx_ctx *ctx, uint32_t n, float *out) {
    if (n == 0) return HCX_SUCCESS;
    k_selftest_log2<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, out);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

size_t packed_smem_bytes(const ChunkArgs &a) {
    if (a.group > kPPp * (uint32_t)kPT) return 0;
    return make_playout(a).total;
}

int launch_packed(hcx_ctx *ctx, const ChunkArgs &a, int kernel) {
    if (a.q.nspec <= 0) return HCX_SUCCESS;
    const size_t sm = packed_smem_bytes(a);
    if (!sm || sm > (size_t)ctx->max_lds)
        return set_error_code(HCX_ERR_INVLD_PARAM, "search: the packed scan does not fit in LDS");
    hipStream_t st = ctx->stream;
    if (kernel == HCX_KERNEL_BATCHED) {
        HCX_HIP(hipFuncSetAttribute((const void *)k_items_packed, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        k_items_packed<<<a.grid, kPT, sm, st>>>(a);
    } else {
        HCX_HIP(hipFuncSetAttribute((const void *)k_spec_packed, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        k_spec_packed<<<(unsigned)a.q.nspec, kPT, sm, st>>>(a);
    }
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

}  // namespace hcx
