// k_chunk.hip -- the fast shared-peak scan over a (bin, peptide-chunk) bucketed index.
//
// Index refinement: each length's peptides (sorted by mass) are cut into
// chunks of TC peptides.  Inside a bin the reference iA is ascending in ion
// id, hence in peptide id, so the entries of one (bin, chunk) pair are a
// contiguous run; bA2 holds where chunk c starts inside bin `bin` (chunk 0 at
// bA[bin], chunk nch at bA[bin+1]), at ba2_index(bin, c, nch): bins in groups
// of eight, each group chunk-major, so a probe's two words share a line, a
// peak's neighbouring bins share one or two, and consecutive chunks of one
// spectrum reuse them from L2.  iA itself is the
// reference array, unchanged; the scan reads iAc, its 16-bit chunk-local
// form (peptide mod TC) << 1 | isY, which needs neither a division nor the
// ion offset at query time and halves the bytes read.
//
// Work item = (spectrum, length, chunk).  The item's (peak, bin) probes are
// taken in groups of at most `group` probes; for one group:
//   R   every probe reads its run bounds [bA2[c], bA2[c+1]); the non-empty
//       runs are compacted by a block scan into one flat sequence of the
//       group's entries, and a bitmap marks where each run starts (one
//       64-bit word per 64-entry tile);
//   P1  each wavefront takes a contiguous share of the tiles; lane j of a
//       tile reads flat entry F + j, whose run is the tile's first run plus
//       the run starts at or below bit j (mbcnt), so consecutive lanes read
//       consecutive iAc entries.  Each entry inside the precursor window adds
//       1 to its peptide's 8-bit hit counter (four counters per LDS word,
//       saturating so that no byte carries); the entry that lifts a counter
//       to min_shp sets the peptide's bit in the qualification bitmap;
// then, once per item:
//   Q   per 32-peptide bitmap word, the number of qualifying peptides before
//       it: a qualifying peptide's slot is its rank, so slots are in peptide
//       order;
//   P2  the entries are read again (from the LDS hit list, past it from iAc;
//       a one-group item keeps its run table) and peptides whose rank falls
//       in the current round of `slots` ranks accumulate exact b/y counts and
//       intensity sums;
//   S   slots are scored: hs = lgf[bc] + lgf[yc] + log10(1+ibc) + log10(1+iyc) - 6
//       (dslim_query.cpp:679-722), histogram + top hit of the spectrum.
// Per-spectrum results from several items merge through global atomics:
// cpsms, histogram bins, and three 64-bit keys whose high half is the
// hyperscore's bit pattern (positive floats order like unsigned ints):
//   top_a = hs | length << 25 | peptide id    (atomicMax)
//   top_b = hs | bc << 16 | yc                (atomicMax)
//   top_n = hs | number of peptides at hs     (CAS merge of (max, count))
// A spectrum whose maximum is shared by more than one peptide goes to the
// in-order heap replay below, as the reference's getMax tie rules require.
//
#include "k_scan_common.h"


namespace hcx {
namespace {


constexpr int kBT = (int)kChunkThreads;   // threads per workgroup of the chunked kernels
constexpr uint32_t kPPT = kProbesPerThread;   // run-table probes per thread
// at most 128 VGPRs: four waves per SIMD (two 512-thread workgroups per CU)
#ifndef HCX_WAVES_PER_SIMD
#define HCX_WAVES_PER_SIMD 4
#endif
constexpr int kWavesPerSimd = HCX_WAVES_PER_SIMD;


// Byte offsets of the dynamic LDS arrays (host and device agree on them).
struct Layout {
    uint32_t map;       // tc/4 words: four 8-bit saturating hit counters each
    uint32_t bmp;       // tc/32 words: qualification bitmap
    uint32_t pre;       // tc/32 words: qualifying peptides before each bitmap word
    uint32_t slot_cnt;  // slots: bc | yc << 16
    uint32_t slot_loc;  // slots u16: local peptide index (chunks hold at most 32768 peptides)
    uint32_t slot_hs;   // slots: hyperscore bits (replay)
    uint32_t slot_sum;  // 2*slots u64 (wide) or u32: ibc, iyc
    uint32_t pk_moz, pk_int;   // max_peaks
    uint32_t cab;       // group + 1: entry index minus flat offset of each non-empty run
    uint32_t cofs;      // group + 1: flat offset of each non-empty run (exclusive prefix)
    uint32_t cint;      // group: intensity of each non-empty run's peak
    uint32_t seg;       // kSegWords u64: run ends of the current segment, a bit per entry
    uint32_t wsum;      // 2 * (kBT / 64) wave partial sums + 2 totals
    uint32_t hist;      // min(hbins, kHistLds): the low bins of the spectrum's histogram
    uint32_t lgf;       // max_shdpeaks + 1 doubles
    uint32_t wlo, whi;  // nlen: windows of the current spectrum
    uint32_t lfirst;    // nlen + 1: first work item of each length in the spectrum (batched)
    uint32_t hl;        // hl_cap u16: hit list of the current item
    uint32_t ctl;       // 8 words: [0..1] control, [2..5] spectrum merge, [6] grabbed item
    uint32_t imax;      // kMaxGrab words: best score of each item of the current grab
    uint32_t red64;     // 2 u64: spectrum merge (entries, probes)
    uint32_t heap;      // replay only: 5 x 256 words, the reference heap (hs, psid, ixx, bc, yc)
    uint32_t total;
};

__host__ __device__ inline uint32_t al16(uint32_t x) { return (x + 15) & ~15u; }

__host__ __device__ inline Layout make_layout(const ChunkArgs &a) {
    Layout l;
    uint32_t off = 0;
    auto take = [&](uint32_t bytes) {
        uint32_t o = off;
        off = al16(off + bytes);
        return o;
    };
    l.map = take(a.tc);
    l.bmp = take(a.tc / 8);
    l.pre = take(a.tc / 8);
    l.slot_cnt = take(4u * a.slots);
    l.slot_loc = take(2u * a.slots);
    l.slot_hs = take(a.replay ? 4u * a.slots : 0u);
    l.slot_sum = take((a.wide ? 16u : 8u) * a.slots);
    l.pk_moz = take(4u * a.max_peaks);
    l.pk_int = take(4u * a.max_peaks);
    l.cab = take(4u * (a.group + 1));
    l.cofs = take(4u * (a.group + 1));
    l.cint = take(4u * a.group);
    l.seg = take(8u * kSegWords);
    l.wsum = take(4u * (2 * (kBT / 64) + 2));
    l.hist = take(4u * min(a.q.hbins, kHistLds));
    l.lgf = take(8u * (a.q.max_shdpeaks + 1));
    l.wlo = take(4u * a.q.nlen);
    l.whi = take(4u * a.q.nlen);
    l.lfirst = take(4u * (a.q.nlen + 1));
    l.hl = take(2u * a.hl_cap);
    l.ctl = take(32u);
    l.imax = take(4u * kMaxGrab);
    l.red64 = take(16u);
    l.heap = take(a.replay ? 5u * 256u * 4u : 0u);
    l.total = off;
    return l;
}

struct Smem {
    lu32 *map, *bmp, *pre, *slot_cnt, *slot_hs, *pk_moz, *pk_int, *cab, *cofs, *cint, *wsum, *hist;
    lu32 *seg32;
    lu64 *seg64;
    lu16 *slot_loc, *hl;
    lu32 *ctl, *heap, *imax, *lfirst;
    lu64 *red64;
    lu64 *slot_sum64;
    lu32 *slot_sum32;
    lf64 *lgf;
    li32 *wlo, *whi;
};

__device__ __forceinline__ Smem bind(unsigned char *raw, const Layout &l) {
    AS_LDS unsigned char *b = (AS_LDS unsigned char *)raw;
    Smem s;
    s.map = (lu32 *)(b + l.map);
    s.bmp = (lu32 *)(b + l.bmp);
    s.pre = (lu32 *)(b + l.pre);
    s.slot_cnt = (lu32 *)(b + l.slot_cnt);
    s.slot_loc = (lu16 *)(b + l.slot_loc);
    s.slot_hs = (lu32 *)(b + l.slot_hs);
    s.slot_sum64 = (lu64 *)(b + l.slot_sum);
    s.slot_sum32 = (lu32 *)(b + l.slot_sum);
    s.pk_moz = (lu32 *)(b + l.pk_moz);
    s.pk_int = (lu32 *)(b + l.pk_int);
    s.cab = (lu32 *)(b + l.cab);
    s.cofs = (lu32 *)(b + l.cofs);
    s.cint = (lu32 *)(b + l.cint);
    s.seg32 = (lu32 *)(b + l.seg);
    s.seg64 = (lu64 *)(b + l.seg);
    s.wsum = (lu32 *)(b + l.wsum);
    s.hist = (lu32 *)(b + l.hist);
    s.lgf = (lf64 *)(b + l.lgf);
    s.wlo = (li32 *)(b + l.wlo);
    s.whi = (li32 *)(b + l.whi);
    s.lfirst = (lu32 *)(b + l.lfirst);
    s.hl = (lu16 *)(b + l.hl);
    s.ctl = (lu32 *)(b + l.ctl);
    s.imax = (lu32 *)(b + l.imax);
    s.red64 = (lu64 *)(b + l.red64);
    s.heap = (lu32 *)(b + l.heap);
    return s;
}


// Run table of the probes [g0, g1) of one item (at most kPPT * kBT probes):
// each (peak, bin) probe's run [bA2[c], bA2[c+1]) in chunk c; the non-empty
// runs are compacted, in probe order, into cab (entry index minus flat
// offset: 32-bit wrap-around, the sum with a flat index is exact; cab[nc]
// = 0), cofs (flat offset, cofs[nc] = E) and cint (the probe's peak
// intensity), and the run ends (last entries) of segment 0 go into the
// bitmap.  Returns (E, nc), uniform.  Every reader of the bitmap and of the
// run table must be past a barrier.
__device__ __forceinline__ uint2 setup_runs(const ChunkArgs &a, const Smem &sm, const LenDesc &L, uint32_t c,
                                            uint32_t g0, uint32_t g1, bool count_probes, Acc &acc) {
    const uint32_t W = 2 * a.q.dF + 1;
    const uint32_t n = g1 - g0;
    uint32_t rs[kPPT], len[kPPT], pk[kPPT];
    uint32_t nr = 0, ne = 0;
#pragma unroll
    for (uint32_t h = 0; h < kPPT; h++) {
        rs[h] = len[h] = pk[h] = 0;
        const uint32_t lp = h * kBT + threadIdx.x;   // any fixed order: P1 and P2 share it
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
    for (uint32_t w = threadIdx.x; w < 2 * kSegWords; w += kBT) sm.seg32[w] = 0;
    uint2 tot;
    const uint2 pre = block_prefix2<kBT>(sm.wsum, nr, ne, &tot);
    uint32_t ci = pre.x, fo = pre.y;
#pragma unroll
    for (uint32_t h = 0; h < kPPT; h++) {
        if (len[h]) {
            sm.cab[ci] = rs[h] - fo;
            sm.cofs[ci] = fo;
            sm.cint[ci] = sm.pk_int[pk[h]];
            fo += len[h];
            if (fo - 1 < kSegEntries) lds_or(&sm.seg32[(fo - 1) >> 5], 1u << ((fo - 1) & 31u));
            ci++;
        }
    }
    if (threadIdx.x == 0) {
        sm.cofs[tot.x] = tot.y;
        sm.cab[tot.x] = 0;
    }
    __syncthreads();
    return make_uint2(ufl(tot.y), ufl(tot.x));
}



// ---- the reference top-K heap (minheap.h:108-177, 258-277), in LDS ----
struct Cell {
    float hs;
    int32_t psid;
    uint32_t ixx, bc, yc;
};
// structure of arrays in LDS: hs (float bits), psid, ixx, bc, yc
struct Heap {
    lu32 *hs, *psid, *ixx, *bc, *yc;
};
__device__ __forceinline__ void heap_put(const Heap &h, int i, const Cell &e) {
    h.hs[i] = __float_as_uint(e.hs);
    h.psid[i] = (uint32_t)e.psid;
    h.ixx[i] = e.ixx;
    h.bc[i] = e.bc;
    h.yc[i] = e.yc;
}
__device__ __forceinline__ Cell heap_get(const Heap &h, int i) {
    Cell e;
    e.hs = __uint_as_float(h.hs[i]);
    e.psid = (int32_t)h.psid[i];
    e.ixx = h.ixx[i];
    e.bc = h.bc[i];
    e.yc = h.yc[i];
    return e;
}
__device__ __forceinline__ float heap_hs(const Heap &h, int i) { return __uint_as_float(h.hs[i]); }

__device__ __forceinline__ void heap_heapify(const Heap &h, int size, int pos) {
    for (;;) {
        const int r = (pos + 1) * 2, l = r - 1;
        int s = pos;
        float sv = heap_hs(h, s);
        if (l < size && heap_hs(h, l) < sv) { s = l; sv = heap_hs(h, l); }
        if (r < size && heap_hs(h, r) < sv) { s = r; sv = heap_hs(h, r); }
        if (s == pos) return;
        const Cell t = heap_get(h, s);
        heap_put(h, s, heap_get(h, pos));
        heap_put(h, pos, t);
        pos = s;
    }
}
__device__ __forceinline__ void heap_insert(const Heap &h, int *size, int cap, const Cell &e) {
    if (*size == cap) {                 // increase_key(0, e)
        if (e.hs < heap_hs(h, 0)) return;
        heap_put(h, 0, e);
        heap_heapify(h, *size, 0);
        return;
    }
    heap_put(h, (*size)++, e);
    for (int i = (*size - 1) / 2; i >= 0; i--) heap_heapify(h, *size, i);
}
__device__ __forceinline__ Cell heap_getmax(const Heap &h, int size) {
    if (size <= 1) return heap_get(h, 0);
    Cell m = heap_get(h, size - 1);
    for (int k = size - 2; k >= size / 2; k--)
        if (heap_hs(h, k) > m.hs) m = heap_get(h, k);
    return m;
}

// Process one (spectrum, length, chunk) item; peaks are already in LDS.
// ORDERED: lane 0 inserts every scored peptide into the heap in peptide order
// (slots are ranks in the qualification bitmap, so slot order is peptide order).
// WIDE: 64-bit intensity sums (otherwise 32-bit, chosen by the host when the
// batch's largest intensity times the most hits a peptide can take fits).
template <bool ORDERED, bool WIDE>
__device__ __forceinline__ void process_item(const ChunkArgs &a, const Smem &sm, lu32 *s_ctl, uint32_t spec,
                                             lu32 *imax_out, uint32_t P, int ixx,
                                             uint32_t c, int32_t lo, int32_t hi, Acc &acc, const Heap *heap = nullptr,
                                             int *hsize = nullptr) {
    const LenDesc L = load_len_u(a, (uint32_t)ixx);
    const uint32_t TC = a.tc;
    const uint32_t base = c * TC;
    const uint32_t ilo = max((uint32_t)lo, base);
    const uint32_t ihi = min((uint32_t)hi, base + TC - 1);
    const uint32_t wlo = ilo - base, wspan = ihi - ilo;   // window in chunk-local ids
    const bool first = (uint32_t)lo / TC == c;
    const uint32_t W = 2 * a.q.dF + 1;
    const uint32_t NP = P * W;
    const uint32_t GP = a.group;
    const uint32_t ng = (NP + GP - 1) / GP;
    // iAc through a buffer resource: 32-bit offsets, and reads past the
    // array return 0 instead of faulting
    const __amdgpu_buffer_rsrc_t rs_iac =
        __builtin_amdgcn_make_buffer_rsrc((void *)L.iAc, 0, (int)(L.n * L.speclen * 2u), kBufDword3);
    const uint32_t lane = threadIdx.x & 63u;
    // the count at which a peptide qualifies (a peptide without hits scores
    // -6 and is never a candidate, so min_shp 0 behaves as 1)
    const uint32_t thr1 = max(a.q.min_shp, 1u) - 1u;

    // P1: 8-bit hit counters.  A counter that reaches 0x80 is pulled back to
    // 0x40: only "hits >= min_shp" (min_shp <= kMaxChunkShp) matters here, and
    // the byte never carries into its neighbour unless more than 0x7F other
    // increments of the same peptide land between one lane's add and its
    // pull-back; a byte seen at >= 0xC0 (halfway there) flags the spectrum.
    // Exactly one increment of a qualifying peptide returns min_shp - 1 (the
    // pull-back never takes a byte below 0x40 > min_shp - 1): that lane sets
    // the peptide's qualification bit.  Entries are chunk-local: local << 1 |
    // isY.  The first hl_cap flat entries of the item (in group order) are
    // also copied to the LDS hit list, so that P2 need not read them again.
    // An item whose window covers its whole chunk (FULL) skips the window
    // test and counts its entries in bulk.
    const bool full = wlo == 0 && wspan == TC - 1;
    uint32_t E = 0, nc = 0, hl_n = 0, seg_cur = 0, nia_w = 0;
    auto p1 = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        for (uint32_t g = 0; g < ng; g++) {
            const uint32_t g0 = g * GP, g1 = min(NP, g0 + GP);
            const uint2 en = setup_runs(a, sm, L, c, g0, g1, first, acc);
            E = en.x;
            nc = en.y;
            seg_cur = 0;
            if (FULL && threadIdx.x == 0) nia_w += E;   // per wavefront otherwise
            const uint32_t lim = min(E, a.hl_cap - hl_n);
            lu16 *hl = sm.hl + hl_n;
            for (uint32_t s0 = 0; s0 < E; s0 += kSegEntries) {
                if (s0 != seg_cur) {
                    build_seg<kBT>(sm.seg32, sm.cofs, nc, s0);
                    seg_cur = s0;
                }
                walk_tiles<kBT, kTileU>(
                    sm.seg64, sm.cofs, nc, s0, min(E, s0 + kSegEntries),
                    [&](const Tile (&t)[kTileU], const uint32_t (&x)[kTileU], const bool (&)[kTileU],
                        uint32_t (&v)[kTileU], uint32_t) {
                        uint32_t gi[kTileU];
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) gi[u] = tile_at(sm.cab, t[u]) + x[u];
                        // lanes outside the share read some entry or, past
                        // the array, 0 (buffer range check); they add 0
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) v[u] = iac_load(rs_iac, gi[u]);
                    },
                    [&](const Tile (&)[kTileU], const uint32_t (&v)[kTileU], const uint32_t (&x)[kTileU],
                        const bool (&ok)[kTileU], uint32_t fs) {
                        uint32_t old[kTileU];
                        bool win[kTileU];
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) {
                            const uint32_t local = v[u] >> 1;
                            win[u] = FULL ? ok[u] : ok[u] && local - wlo <= wspan;
                            if (!FULL) nia_w += (uint32_t)__popcll(__ballot(win[u]));
                            // every lane adds: 0 outside the window or the share
                            old[u] = lds_add_rtn(&sm.map[local >> 2], win[u] ? 1u << ((v[u] & 6u) << 2) : 0u);
                        }
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) {
                            const uint32_t local = v[u] >> 1;
                            const uint32_t sh = (v[u] & 6u) << 2;
                            const uint32_t ob = (old[u] >> sh) & 0xFFu;
                            if (win[u] && (ob == thr1 || ob >= 0x7Fu)) {
                                if (ob == thr1) lds_or(&sm.bmp[local >> 5], 1u << (local & 31u));
                                else if (ob == 0x7Fu) lds_add(&sm.map[local >> 2], 0u - (0x40u << sh));
                                else if (ob >= 0xC0u) acc.flags |= HCX_RES_COUNT_OVF;
                            }
                        }
                        if (fs < lim) {
#pragma unroll
                            for (uint32_t u = 0; u < kTileU; u++)
                                if (ok[u] && x[u] < lim) hl[x[u]] = (uint16_t)v[u];
                        }
                    });
            }
            hl_n += lim;
            __syncthreads();
        }
    };
    if (full) p1(std::true_type{});
    else p1(std::false_type{});
    if (lane == 0) acc.nia += nia_w;

    // Q: qualifying peptides before each bitmap word; the counters are done
    // with and are cleared for the next item
    {
        const uint32_t nbw = TC >> 5;
        const uint32_t per = (nbw + kBT - 1) / kBT;
        const uint32_t t0 = min(nbw, threadIdx.x * per);
        const uint32_t t1 = min(nbw, t0 + per);
        uint32_t tsum = 0;
        for (uint32_t bw = t0; bw < t1; bw++) tsum += __popc(sm.bmp[bw]);
        for (uint32_t i = threadIdx.x; i < TC / 16; i += kBT) ((lu32x4 *)sm.map)[i] = (u32x4){0u, 0u, 0u, 0u};
        uint32_t run = block_prefix<kBT>(sm.wsum, tsum);
        for (uint32_t bw = t0; bw < t1; bw++) {
            const uint32_t cnt = __popc(sm.bmp[bw]);
            sm.pre[bw] = run;
            run += cnt;
        }
        if (threadIdx.x == kBT - 1) s_ctl[0] = run;
        __syncthreads();
    }
    const uint32_t Qn = ufl((uint32_t)s_ctl[0]);
    float imax = 0.0f;   // this thread's best score in the item (batched item list)

    for (uint32_t r0 = 0; r0 < Qn; r0 += a.slots) {
        const uint32_t ns = min(a.slots, Qn - r0);
        // the slots start at zero: S clears what it has read, except in the
        // replay, whose heap feed reads them after S
        if (ORDERED) {
            for (uint32_t sl = threadIdx.x; sl < ns; sl += kBT) {
                sm.slot_cnt[sl] = 0;
                if (WIDE) {
                    sm.slot_sum64[2 * sl] = 0;
                    sm.slot_sum64[2 * sl + 1] = 0;
                } else {
                    sm.slot_sum32[2 * sl] = 0;
                    sm.slot_sum32[2 * sl + 1] = 0;
                }
            }
            __syncthreads();
        }
        // P2: exact b/y counts and intensity sums of slots [r0, r0 + ns),
        // group by group: the listed entries from the hit list, the rest
        // from the entry array again.  Only the bitmap matters here: a
        // peptide outside the window never has its bit set.
        uint32_t hb = 0;
        for (uint32_t g = 0; g < ng; g++) {
            const uint32_t g0 = g * GP, g1 = min(NP, g0 + GP);
            if (ng > 1) {
                const uint2 en = setup_runs(a, sm, L, c, g0, g1, false, acc);
                E = en.x;
                nc = en.y;
                seg_cur = 0;
            }
            const uint32_t lim = min(E, a.hl_cap - hb);
            const lu16 *hl = sm.hl + hb;
            for (uint32_t s0 = 0; s0 < E; s0 += kSegEntries) {
                if (s0 != seg_cur) {
                    build_seg<kBT>(sm.seg32, sm.cofs, nc, s0);
                    seg_cur = s0;
                }
                walk_tiles<kBT, kTileU>(
                    sm.seg64, sm.cofs, nc, s0, min(E, s0 + kSegEntries),
                    [&](const Tile (&t)[kTileU], const uint32_t (&x)[kTileU], const bool (&ok)[kTileU],
                        uint32_t (&v)[kTileU], uint32_t fs) {
                        if (fs + 64 * kTileU <= lim) {   // the whole step is in the hit list
#pragma unroll
                            for (uint32_t u = 0; u < kTileU; u++) v[u] = hl[x[u]];
                            return;
                        }
                        uint32_t gi[kTileU];
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) gi[u] = tile_at(sm.cab, t[u]) + x[u];
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) v[u] = iac_load(rs_iac, gi[u]);
                        if (fs < lim) {
#pragma unroll
                            for (uint32_t u = 0; u < kTileU; u++)
                                if (x[u] < lim) v[u] = hl[x[u]];
                        }
                    },
                    [&](const Tile (&t)[kTileU], const uint32_t (&v)[kTileU], const uint32_t (&)[kTileU],
                        const bool (&ok)[kTileU], uint32_t) {
                        uint32_t bits[kTileU];
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) bits[u] = sm.bmp[v[u] >> 6];
#pragma unroll
                        for (uint32_t u = 0; u < kTileU; u++) {
                            const uint32_t bit = (v[u] >> 1) & 31u;
                            if (!(ok[u] && ((bits[u] >> bit) & 1u))) continue;
                            const uint32_t local = v[u] >> 1;
                            const uint32_t sl = sm.pre[local >> 5] + __popc(bits[u] & ((1u << bit) - 1u)) - r0;
                            if (sl >= ns) continue;
                            const uint32_t isY = v[u] & 1u;
                            const uint32_t it = tile_at(sm.cint, t[u]);
                            lds_add(&sm.slot_cnt[sl], isY ? 0x10000u : 1u);
                            if (WIDE) lds_add64(&sm.slot_sum64[2 * sl + isY], (unsigned long long)it);
                            else lds_add(&sm.slot_sum32[2 * sl + isY], it);
                            sm.slot_loc[sl] = (uint16_t)local;
                        }
                    });
            }
            hb += lim;
            if (ng > 1) __syncthreads();
        }
        __syncthreads();
        // S: score the slots
        for (uint32_t sl = threadIdx.x; sl < ns; sl += kBT) {
            const uint32_t loc = sm.slot_loc[sl];
            const uint32_t cw = sm.slot_cnt[sl];
            const uint32_t bc = cw & 0xFFFFu, yc = cw >> 16;
            const unsigned long long ib = WIDE ? sm.slot_sum64[2 * sl] : sm.slot_sum32[2 * sl];
            const unsigned long long iy = WIDE ? sm.slot_sum64[2 * sl + 1] : sm.slot_sum32[2 * sl + 1];
            if (ORDERED) {
                sm.slot_hs[sl] = 0u;   // +0.0f: not inserted
            } else {
                sm.slot_cnt[sl] = 0;
                if (WIDE) {
                    sm.slot_sum64[2 * sl] = 0;
                    sm.slot_sum64[2 * sl + 1] = 0;
                } else {
                    sm.slot_sum32[2 * sl] = 0;
                    sm.slot_sum32[2 * sl + 1] = 0;
                }
            }
            if ((uint16_t)(bc + yc) < a.q.min_shp) continue;
            uint32_t bi = bc, yi = yc;
            if (bi > a.q.max_shdpeaks) { bi = a.q.max_shdpeaks; acc.flags |= HCX_RES_LGF_CLAMP; }
            if (yi > a.q.max_shdpeaks) { yi = a.q.max_shdpeaks; acc.flags |= HCX_RES_LGF_CLAMP; }
            double t = sm.lgf[bi] + sm.lgf[yi];
            t = t + log10((double)(1ull + ib));
            t = t + log10((double)(1ull + iy));
            t = t - 6.0;
            const float hs = (float)t;
            if (!(hs > 0.0f)) continue;
            if (ORDERED) {
                sm.slot_hs[sl] = __float_as_uint(hs);
                continue;
            }
            acc.cpsms++;
            imax = fmaxf(imax, hs);
            int bin = (int)((double)(hs * 10.0f) + 0.5);
            if (bin >= (int)a.q.hbins) { bin = (int)a.q.hbins - 1; acc.flags |= HCX_RES_HIST_CLAMP; }
            if ((uint32_t)bin < kHistLds) lds_add(&sm.hist[bin], 1u);
            else glb_add(a.q.hist + (size_t)spec * a.q.hbins + bin, 1u);
            if (hs > acc.best_hs) {
                acc.best_hs = hs;
                acc.best_cnt = 1;
                acc.best_id = ((uint32_t)ixx << 25) | (base + loc);
                acc.best_by = (bc << 16) | yc;
            } else if (hs == acc.best_hs) {
                acc.best_cnt++;
            }
        }
        // the item's best score, for the replay (positive floats order as uints)
        // (one LDS atomic per wavefront: same-address atomics serialise)
        if (!ORDERED && imax_out) {
            const uint32_t wm = wave_max_last(__float_as_uint(imax));
            if ((threadIdx.x & 63u) == 63u && wm) lds_max(imax_out, wm);
        }
        __syncthreads();
        if (ORDERED) {
            // wave 0 feeds the round's scored slots to the heap in slot order;
            // once the heap is full, a slot below the root would be rejected
            // (increase_key, minheap.h:218-225), and the root only grows, so a
            // ballot skips every such slot up to the next one that is taken
            if (threadIdx.x < 64) {
                const int cap = a.q.heap_cap;
                for (uint32_t b0 = 0; b0 < ns; b0 += 64) {
                    const uint32_t sl = b0 + threadIdx.x;
                    const float hs = sl < ns ? __uint_as_float(sm.slot_hs[sl]) : 0.0f;
                    unsigned long long mask = __ballot(hs > 0.0f);
                    while (mask) {
                        const int hsz = __shfl(*hsize, 0, 64);
                        const float root = hsz == cap ? heap_hs(*heap, 0) : -1.0f;
                        const unsigned long long take = mask & __ballot(hs >= root);
                        if (!take) break;
                        const int first = __ffsll((long long)take) - 1;
                        if (threadIdx.x == 0) {
                            const uint32_t s2 = b0 + (uint32_t)first;
                            const uint32_t cw = sm.slot_cnt[s2];
                            const Cell e = {__uint_as_float(sm.slot_hs[s2]), (int32_t)(base + sm.slot_loc[s2]),
                                            (uint32_t)ixx, cw & 0xFFFFu, cw >> 16};
                            heap_insert(*heap, hsize, cap, e);
                        }
                        mask &= ~((2ull << first) - 1ull);   // slots up to `first` are done
                    }
                }
            }
            __syncthreads();
        }
    }
    // the qualification bitmap's last reader (P2) is past a barrier; with no
    // qualifying peptide it is already clear
    if (Qn)
        for (uint32_t bw = threadIdx.x; bw < (TC >> 5); bw += kBT) sm.bmp[bw] = 0;
}


__device__ __forceinline__ void init_smem(const ChunkArgs &a, const Smem &sm) {
    for (uint32_t i = threadIdx.x; i <= a.q.max_shdpeaks; i += kBT) sm.lgf[i] = a.q.lgf[i];
    for (uint32_t i = threadIdx.x; i < a.tc / 4; i += kBT) sm.map[i] = 0;
    for (uint32_t i = threadIdx.x; i < a.tc / 32; i += kBT) sm.bmp[i] = 0;
    if (threadIdx.x < 8) sm.ctl[threadIdx.x] = 0;
    for (uint32_t i = threadIdx.x; i < a.slots; i += kBT) {
        sm.slot_cnt[i] = 0;
        if (a.wide) {
            sm.slot_sum64[2 * i] = 0;
            sm.slot_sum64[2 * i + 1] = 0;
        } else {
            sm.slot_sum32[2 * i] = 0;
            sm.slot_sum32[2 * i + 1] = 0;
        }
    }
    for (uint32_t i = threadIdx.x; i < min(a.q.hbins, kHistLds); i += kBT) sm.hist[i] = 0;
}

// Two-pass form, one spectrum at a time: with a.fb_list, the listed spectra
// (those the packed kernel leaves), taken in turn by a persistent grid;
// otherwise one workgroup per spectrum.  Items in order.
template <bool WIDE>
__global__ __launch_bounds__(kBT, kWavesPerSimd) void k_spec_chunked(ChunkArgs a) {
    // all LDS is dynamic, so the hit counters start at address 0
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const Smem sm = bind(smem_raw, make_layout(a));
    lu32 *s_ctl = sm.ctl, *s_red = sm.ctl + 2;
    lu64 *s_red64 = sm.red64;
    init_smem(a, sm);
    if (threadIdx.x == 0) {
        s_ctl[0] = 0;
    }
    const uint32_t nwork = a.fb_list ? ufl(*(const AS_GLB uint32_t *)a.fb_count) : (uint32_t)a.q.nspec;
    for (uint32_t wi = blockIdx.x; wi < nwork; wi += a.fb_list ? gridDim.x : nwork) {
        const uint32_t s = a.fb_list ? ufl(((const AS_GLB uint32_t *)a.fb_list)[wi]) : wi;
        __syncthreads();
        const uint32_t P = load_peaks<kBT>(a, sm.pk_moz, sm.pk_int, s);
        for (int l = threadIdx.x; l < a.q.nlen; l += kBT) {
            const int2 w = a.win[(size_t)s * a.q.nlen + l];
            sm.wlo[l] = w.x;
            sm.whi[l] = w.y;
        }
        __syncthreads();
        Acc acc;
        acc_reset(acc);
        for (int ixx = 0; ixx < a.q.nlen; ixx++) {
            const int32_t lo = ufl((int32_t)sm.wlo[ixx]), hi = ufl((int32_t)sm.whi[ixx]);
            if (lo > hi) continue;
            for (uint32_t c = (uint32_t)lo / a.tc; c <= (uint32_t)hi / a.tc; c++)
                process_item<false, WIDE>(a, sm, s_ctl, s, nullptr, P, ixx, c, lo, hi, acc);
        }
        flush_spectrum<kBT>(a, sm.hist, s, acc, s_red, s_red64);
    }
}

// Batched two-pass form: persistent workgroups pull (spectrum, length, chunk)
// items (run_items).
template <bool WIDE>
__global__ __launch_bounds__(kBT, kWavesPerSimd) void k_items_chunked(ChunkArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const Smem sm = bind(smem_raw, make_layout(a));
    lu32 *s_ctl = sm.ctl, *s_red = sm.ctl + 2, *s_item = sm.ctl + 6;
    lu64 *s_red64 = sm.red64;
    init_smem(a, sm);
    if (threadIdx.x == 0) {
        s_ctl[0] = 0;
    }
    Acc acc;
    acc_reset(acc);
    uint32_t P = 0;
    run_items<kBT>(
        a, s_item, sm.imax, sm.wlo, sm.lfirst,
        [&](uint32_t s) {
            P = load_peaks<kBT>(a, sm.pk_moz, sm.pk_int, s);
            load_windows<kBT>(a, s, sm.wlo, sm.whi, sm.lfirst);
        },
        [&](uint32_t s, uint32_t l, uint32_t c, lu32 *imax_out) {
            process_item<false, WIDE>(a, sm, s_ctl, s, imax_out, P, (int)l, c, ufl((int32_t)sm.wlo[l]),
                                      ufl((int32_t)sm.whi[l]), acc);
        },
        [&](uint32_t s) { flush_spectrum<kBT>(a, sm.hist, s, acc, s_red, s_red64); });
}

// Tie replay: spectra whose maximum hyperscore is shared by several peptides.
// One workgroup per listed spectrum runs its items in the reference order
// (length, then peptide id) and inserts every scored peptide into the heap,
// whose getMax() then picks the top hit exactly as the reference does.
template <bool WIDE>
__global__ __launch_bounds__(kBT) void k_replay_chunked(ChunkArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const Smem sm = bind(smem_raw, make_layout(a));
    lu32 *s_ctl = sm.ctl;
    const Heap heap = {sm.heap, sm.heap + 256, sm.heap + 512, sm.heap + 768, sm.heap + 1024};
    init_smem(a, sm);
    if (threadIdx.x == 0) {
        s_ctl[0] = 0;
    }
    const uint32_t nwork = ((const AS_GLB uint32_t *)a.q.tie)[0];
    // tied spectra are taken one at a time from a counter after the list
    // (zeroed before the scan): their replays differ widely in length
    __shared__ uint32_t s_grab;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_grab = atomicAdd(&a.q.tie[a.q.nspec + 1], 1u);
        __syncthreads();
        const uint32_t wi = s_grab;
        if (wi >= nwork) break;
        const uint32_t s = ((const AS_GLB uint32_t *)a.q.tie)[1 + wi];
        __syncthreads();
        const uint32_t P = load_peaks<kBT>(a, sm.pk_moz, sm.pk_int, s);
        for (int l = threadIdx.x; l < a.q.nlen; l += kBT) {
            const int2 w = a.win[(size_t)s * a.q.nlen + l];
            sm.wlo[l] = w.x;
            sm.whi[l] = w.y;
        }
        // the heap's size and root as thread 0 last published them.  A
        // spectrum the batched kernel scored has its items numbered from
        // item_base[s] in this loop's order (a spectrum scored elsewhere
        // has none), and item_max holds their best scores.
        uint32_t ir = 0;
        bool skip_ok = false;
        if (a.item_max != nullptr && a.item_base != nullptr) {
            ir = ufl(((const AS_GLB uint32_t *)a.item_base)[s]);
            const uint32_t ie = ufl(((const AS_GLB uint32_t *)a.item_base)[s + 1]);
            skip_ok = ie > ir && ie <= a.item_cap;
        }
        if (threadIdx.x == 0) {
            sm.ctl[6] = 0;
            sm.ctl[7] = 0;
        }
        __syncthreads();
        int hsize = 0;   // meaningful on lane 0 only
        Acc acc;
        acc_reset(acc);
        for (int ixx = 0; ixx < a.q.nlen; ixx++) {
            const int32_t lo = ufl((int32_t)sm.wlo[ixx]), hi = ufl((int32_t)sm.whi[ixx]);
            if (lo > hi) continue;
            for (uint32_t c = (uint32_t)lo / a.tc; c <= (uint32_t)hi / a.tc; c++) {
                if (skip_ok) {
                    // every candidate of an item whose best score is below the
                    // full heap's root is rejected (minheap.h:111-114): skip it
                    const uint32_t m = ufl(((const AS_GLB uint32_t *)a.item_max)[ir++]);
                    const uint32_t hsz = ufl((uint32_t)sm.ctl[6]), rootb = ufl((uint32_t)sm.ctl[7]);
                    if (m == 0u || (hsz == (uint32_t)a.q.heap_cap && __uint_as_float(m) < __uint_as_float(rootb)))
                        continue;
                }
                process_item<true, WIDE>(a, sm, s_ctl, s, nullptr, P, ixx, c, lo, hi, acc, &heap, &hsize);
                if (skip_ok) {
                    if (threadIdx.x == 0) {
                        sm.ctl[6] = (uint32_t)hsize;
                        sm.ctl[7] = __float_as_uint(heap_hs(heap, 0));
                    }
                    __syncthreads();
                }
            }
        }
        if (threadIdx.x == 0 && hsize > 0) {
            const Cell m = heap_getmax(heap, hsize);
            hcx_result r = a.q.res[s];
            r.hyperscore = m.hs;
            r.psid = m.psid;
            r.idxoffset = (uint16_t)m.ixx;
            r.bc = m.bc;
            r.yc = m.yc;
            r.sharedions = (uint16_t)(m.bc + m.yc);
            r.totalions = (uint16_t)load_len(a, (int)m.ixx).speclen;
            r.flags |= HCX_RES_TIE_REPLAYED;
            a.q.res[s] = r;
        }
    }
}


// Per spectrum (one wavefront): the precursor window of every length
// (dslim_query.cpp:854-916), whether the packed kernel scores it, and -- for
// the batched forms -- its number of (length, chunk) work items.  Spectra the
// packed kernel cannot score exactly go to the two-pass kernel's list.
__global__ __launch_bounds__(256) void k_windows(ChunkArgs a, WinArgs w) {
    const uint32_t s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (s >= (uint32_t)a.q.nspec) return;
    const uint32_t D = a.q.max_mass + 2;
    const float prec = a.q.prec[s];
    uint32_t cnt = 0;
    for (int l = (int)lane; l < a.q.nlen; l += 64) {
        int32_t lo, hi;
        window(a.q.lens[l], prec, a.q.dM, D, &lo, &hi);
        w.win[(size_t)s * a.q.nlen + l] = make_int2(lo, hi);
        if (lo <= hi) cnt += (uint32_t)hi / a.tc - (uint32_t)lo / a.tc + 1;
    }
    cnt = (uint32_t)__builtin_amdgcn_readlane((int)wave_reduce_last(cnt, [](uint32_t u, uint32_t v) { return u + v; }),
                                              63);
    if (lane != 0) return;
    // a series takes at most half hits per peak window of 2 * dF + 1 bins
    // (one ion per entry, the hits of one ion from one window): its count
    // must stay below 256 and its intensity sum below 2^24
    const bool packed = w.pack && (uint64_t)w.half * w.mult[s] <= 255u &&
                        (uint64_t)w.half * w.wsum[s] < (1ull << 24);
    w.packed[s] = packed ? 1 : 0;
    if (!packed) w.acc[s].flags = HCX_RES_TWO_PASS;   // before any scan kernel ORs its flags in
    if (w.batched) w.counts[s] = (packed || !w.pack) ? cnt : 0u;
    if (w.pack && !packed && cnt) w.fb_list[atomicAdd(w.fb_count, 1u)] = s;
}

// item_spec[i] = the spectrum holding item 64 * i (one thread per spectrum).
__global__ void k_item_spec(const uint32_t *item_base, int32_t nspec, uint32_t *item_spec) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (uint32_t)nspec) return;
    const uint32_t b0 = item_base[s], b1 = item_base[s + 1];
    for (uint32_t i = (b0 + 63) >> 6; (i << 6) < b1; i++) item_spec[i] = s;
}

// SpecAcc -> hcx_result; tied maxima go to the replay list.
__global__ void k_finalize(ChunkArgs a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (uint32_t)a.q.nspec) return;
    const SpecAcc g = a.acc[s];
    hcx_result r;
    r.hyperscore = 0.0f;
    r.psid = 0;
    r.idxoffset = 0;
    r.sharedions = 0;
    r.totalions = 0;
    r.bc = r.yc = 0;
    r.cpsms = g.cpsms;
    r.flags = (uint16_t)g.flags;
    r.maxhypscore = r.mu = r.beta = r.minhypscore = r.nexthypscore = 0;
    r.e_x = 0.0;
    const uint32_t cnt = (uint32_t)g.top_n;
    // a spectrum whose 8-bit counters neared overflow may have a wrong
    // candidate set: it is re-scored exactly by the cursor kernel (and its
    // ties replayed there), not by the chunked replay
    if ((g.flags & HCX_RES_COUNT_OVF) || (a.force_redo && s % a.force_redo == 0)) {
        a.redo[1 + atomicAdd(&a.redo[0], 1u)] = s;
    } else if (cnt >= 1) {
        r.hyperscore = __uint_as_float((uint32_t)(g.top_n >> 32));
        if (cnt == 1) {
            const uint32_t ixx = (uint32_t)((g.top_a >> 25) & 0x7F);
            r.psid = (int32_t)(g.top_a & kLow25);
            r.idxoffset = (uint16_t)ixx;
            r.bc = (uint32_t)((g.top_b >> 16) & 0xFFFF);
            r.yc = (uint32_t)(g.top_b & 0xFFFF);
            r.sharedions = (uint16_t)(r.bc + r.yc);
            r.totalions = (uint16_t)a.q.lens[ixx].speclen;
        } else {
            const uint32_t slot = atomicAdd(&a.q.tie[0], 1u);
            a.q.tie[1 + slot] = s;
        }
    }
    a.q.res[s] = r;
    if (a.q.stats) {
        a.q.stats[s].n_ia = g.n_ia;
        a.q.stats[s].n_probe = g.n_probe;
    }
}

__global__ void k_chunk_offsets(const uint32_t *iA, const uint32_t *bA, uint32_t NB, uint32_t nch, uint32_t TC,
                                uint32_t speclen, uint32_t *bA2) {
    // one thread per word, in storage order (ba2_index)
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ba2_words(NB, nch)) return;
    const uint64_t gw = (uint64_t)kBa2Group * (nch + 1);   // words of a bin group
    const uint32_t c = (uint32_t)(t % gw / kBa2Group);
    const uint32_t bin = (uint32_t)(t / gw * kBa2Group + t % kBa2Group);
    if (bin >= NB) {   // the last group's padding
        bA2[t] = 0;
        return;
    }
    const uint32_t s0 = bA[bin], e0 = bA[bin + 1];
    uint32_t r;
    if (c == 0) {
        r = s0;
    } else if (c == nch) {
        r = e0;
    } else {
        const uint64_t key = (uint64_t)c * TC * speclen;
        uint32_t lo = s0, hi = e0;
        while (lo < hi) {
            const uint32_t m = lo + (hi - lo) / 2;
            if ((uint64_t)iA[m] < key) lo = m + 1;
            else hi = m;
        }
        r = lo;
    }
    bA2[t] = r;
}

// Chunk-local copy of iA: entry e = (peptide mod TC) << 1 | isY, where the
// ion id iA[e] = peptide * speclen + ion and isY = ion >= speclen / 2.
__global__ void k_chunk_local(const uint32_t *iA, uint64_t n, uint32_t speclen, uint32_t TC, uint16_t *iAc) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t raw = iA[e];
    const uint32_t ppid = raw / speclen;
    const uint32_t isY = (raw - ppid * speclen) >= speclen / 2 ? 1u : 0u;
    iAc[e] = (uint16_t)(((ppid & (TC - 1)) << 1) | isY);
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

size_t chunk_smem_bytes(const ChunkArgs &a) { return make_layout(a).total; }

int launch_chunk_offsets(hcx_ctx *ctx, const uint32_t *iA, const uint32_t *bA, uint32_t NB, uint32_t nch, uint32_t TC,
                         uint32_t speclen, uint32_t *bA2) {
    const uint64_t n = ba2_words(NB, nch);
    k_chunk_offsets<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(iA, bA, NB, nch, TC, speclen, bA2);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_chunk_local(hcx_ctx *ctx, const uint32_t *iA, uint64_t n, uint32_t speclen, uint32_t TC, uint16_t *iAc) {
    if (n == 0) return HCX_SUCCESS;
    k_chunk_local<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(iA, n, speclen, TC, iAc);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_fill_u32(hcx_ctx *ctx, uint32_t *p, uint64_t n, uint32_t v) {
    if (n == 0) return HCX_SUCCESS;
    k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p, n, v);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_windows(hcx_ctx *ctx, const ChunkArgs &a, const WinArgs &w) {
    if (a.q.nspec <= 0) return HCX_SUCCESS;
    k_windows<<<(unsigned)((a.q.nspec + 3) / 4), 256, 0, ctx->stream>>>(a, w);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_item_spec(hcx_ctx *ctx, const uint32_t *item_base, int32_t nspec, uint32_t *item_spec) {
    if (nspec <= 0) return HCX_SUCCESS;
    k_item_spec<<<(unsigned)((nspec + 255) / 256), 256, 0, ctx->stream>>>(item_base, nspec, item_spec);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

// kernel HCX_KERNEL_BATCHED: the persistent item kernel (grid a.grid);
// otherwise the spectrum kernel: persistent over a.fb_list (grid a.grid), or
// one workgroup per spectrum.
template <bool WIDE>
static int launch_chunked_t(hcx_ctx *ctx, const ChunkArgs &a, int kernel) {
    hipStream_t st = ctx->stream;
    const size_t sm = chunk_smem_bytes(a);
    if (kernel == HCX_KERNEL_BATCHED) {
        HCX_HIP(hipFuncSetAttribute((const void *)k_items_chunked<WIDE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)sm));
        k_items_chunked<WIDE><<<a.grid, kBT, sm, st>>>(a);
    } else {
        HCX_HIP(hipFuncSetAttribute((const void *)k_spec_chunked<WIDE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)sm));
        k_spec_chunked<WIDE><<<a.fb_list ? a.grid : (unsigned)a.q.nspec, kBT, sm, st>>>(a);
    }
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_chunked(hcx_ctx *ctx, const ChunkArgs &a, int kernel) {
    if (a.q.nspec <= 0) return HCX_SUCCESS;
    return a.wide ? launch_chunked_t<true>(ctx, a, kernel) : launch_chunked_t<false>(ctx, a, kernel);
}

int launch_finalize(hcx_ctx *ctx, const ChunkArgs &a) {
    if (a.q.nspec <= 0) return HCX_SUCCESS;
    k_finalize<<<(a.q.nspec + 255) / 256, 256, 0, ctx->stream>>>(a);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

template <bool WIDE>
static int launch_replay_t(hcx_ctx *ctx, const ChunkArgs &a) {
    const size_t sm = chunk_smem_bytes(a);
    HCX_HIP(hipFuncSetAttribute((const void *)k_replay_chunked<WIDE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sm));
    const int grid = std::max(1, std::min(a.q.nspec, ctx->num_cu * (int)ctx->blocks_per_cu));
    k_replay_chunked<WIDE><<<grid, kBT, sm, ctx->stream>>>(a);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_replay_chunked(hcx_ctx *ctx, ChunkArgs &a) {
    if (a.q.nspec <= 0) return HCX_SUCCESS;
    ChunkArgs r = a;
    r.replay = 1;   // the replay layout adds the slots' hyperscores
    // its hit list takes what is left of a workgroup's share (the layout
    // holds the heap)
    r.hl_cap = 0;
    const size_t base = chunk_smem_bytes(r);
    const size_t share = (size_t)ctx->max_lds / std::max<uint32_t>(1, ctx->blocks_per_cu);
    const size_t fixed = 16;   // the kernel's static LDS (the work counter's copy)
    if (share > base + fixed + 4096)
        r.hl_cap = std::min<uint32_t>(a.hl_cap, (uint32_t)(((share - base - fixed) / 2) & ~(size_t)63));
    if (chunk_smem_bytes(r) + fixed > (size_t)ctx->max_lds)
        return set_error_code(HCX_ERR_INVLD_PARAM, "search: the replay layout does not fit in LDS");
    return r.wide ? launch_replay_t<true>(ctx, r) : launch_replay_t<false>(ctx, r);
}

}  // namespace hcx
