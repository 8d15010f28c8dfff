// k_scan_common.h -- device helpers shared by the scan kernels (k_chunk.hip,
// the two-pass exact-count scan, and k_packed.hip, the single-pass packed
// scan): address-space-qualified LDS / global access, wavefront scans and
// reductions over DPP, the precursor window, and the tile walk over a flat run
// table.  Workgroup-wide helpers take the workgroup size BT as a template
// argument.
//
// Every LDS pointer is address-space-3 qualified and every global pointer on
// the scan path address-space-1 qualified, so no access compiles to a flat
// instruction (a flat load counts against both the vector-memory and the LDS
// wait counters and would serialise the LDS atomics behind HBM reads).
#pragma once
#include "hcx_internal.h"

#include <type_traits>

namespace hcx {
namespace {

#define AS_LDS __attribute__((address_space(3)))
#define AS_GLB __attribute__((address_space(1)))
typedef AS_LDS uint32_t lu32;
typedef AS_LDS uint16_t lu16;
typedef AS_LDS int32_t li32;
typedef AS_LDS unsigned long long lu64;
typedef AS_LDS double lf64;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef AS_LDS u32x4 lu32x4;
typedef const AS_GLB uint32_t gcu32;
typedef AS_GLB uint32_t gu32;
typedef AS_GLB unsigned long long gu64;

constexpr uint32_t kNoProbe = 0xFFFFFFFFu;
constexpr unsigned long long kLow25 = (1ull << 25) - 1;
// A group's flat entries are walked in segments of kSegEntries; the run
// starts of the current segment are a bitmap of kSegWords 64-bit words, one
// per 64-entry tile.
constexpr uint32_t kSegWords = 256;
constexpr uint32_t kSegEntries = 64u * kSegWords;
#ifndef HCX_TILE_U
#define HCX_TILE_U 8
#endif
constexpr uint32_t kTileU = HCX_TILE_U;   // tiles per wavefront step (loads in flight)
// histogram bins kept in LDS (hyperscores below 51.15); the rare higher
// bins go to the global histogram directly
constexpr uint32_t kHistLds = 512;

__device__ __forceinline__ uint32_t ldg(const uint32_t *p, size_t i) { return ((gcu32 *)p)[i]; }
// raw buffer resource word 3 on gfx9-family parts (gfx950 included): byte
// offsets, range-checked against num_records
constexpr int kBufDword3 = 0x00020000;
__device__ __forceinline__ uint32_t iac_load(__amdgpu_buffer_rsrc_t r, uint32_t entry) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)(entry << 1), 0, 0);
}
__device__ __forceinline__ void lds_add(lu32 *p, uint32_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_add_rtn(lu32 *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add64(lu64 *p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_and(lu32 *p, uint32_t v) {
    __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_max(lu32 *p, uint32_t v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_or(lu32 *p, uint32_t v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void glb_add(T *p, T v) {
    __hip_atomic_fetch_add((AS_GLB T *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void glb_or(uint32_t *p, uint32_t v) {
    __hip_atomic_fetch_or((gu32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void glb_max64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_fetch_max((gu64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t udiv(uint32_t n, uint32_t m, uint32_t s) {
    if (s == 0) return n;
    uint32_t t = __umulhi(m, n);
    return (t + ((n - t) >> 1)) >> (s - 1);
}

__device__ __forceinline__ uint32_t mass_lb(const float *mass, uint32_t lo, uint32_t hi, float v) {
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (mass[m] < v) lo = m + 1;
        else hi = m;
    }
    return lo;
}
__device__ __forceinline__ uint32_t mass_ub(const float *mass, uint32_t lo, uint32_t hi, float v) {
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (v < mass[m]) hi = m;
        else lo = m + 1;
    }
    return lo;
}
__device__ __forceinline__ void dir_range(const uint32_t *mdir, uint32_t D, uint32_t n, float v, uint32_t *lo,
                                          uint32_t *hi) {
    if (!(v > 0.0f)) {
        *lo = 0;
        *hi = mdir[1];
        return;
    }
    uint32_t d = (v >= (float)(D - 1)) ? D - 1 : (uint32_t)v;
    *lo = mdir[d];
    *hi = (d + 1 < D) ? mdir[d + 1] : n;
}
// dslim_query.cpp:854-916 as lower/upper bounds over float masses
__device__ __forceinline__ void window(const LenDesc &L, float prec, double dM, uint32_t D, int32_t *wlo,
                                       int32_t *whi) {
    *wlo = 1;
    *whi = 0;
    if (L.n == 0 || dM < 0.0 || prec != prec) return;
    const float p1 = (float)((double)prec - dM);
    const float p2 = (float)((double)prec + dM);
    uint32_t lo, hi;
    dir_range(L.mdir, D, L.n, p1, &lo, &hi);
    const uint32_t a = mass_lb(L.mass, lo, hi, p1);
    dir_range(L.mdir, D, L.n, p2, &lo, &hi);
    const uint32_t b = mass_ub(L.mass, lo, hi, p2);
    *wlo = (int32_t)a;
    *whi = (int32_t)b - 1;
}

// Values every lane of the workgroup holds alike, moved to scalar registers
// (an LDS or global read lands in a vector register even when uniform).
__device__ __forceinline__ uint32_t ufl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int32_t ufl(int32_t v) { return (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)v); }
template <typename T>
__device__ __forceinline__ T *ufl(T *p) {
    const uint64_t v = (uint64_t)(uintptr_t)p;
    return (T *)(uintptr_t)(((uint64_t)ufl((uint32_t)(v >> 32)) << 32) | ufl((uint32_t)v));
}
struct Acc {
    float best_hs;
    uint32_t best_cnt;
    uint32_t best_id;    // length << 25 | peptide id
    uint32_t best_by;    // bc << 16 | yc
    uint32_t cpsms, flags;
    uint32_t nia, nprobe;   // per thread and spectrum: far below 2^32
};

__device__ __forceinline__ void acc_reset(Acc &a) {
    a.best_hs = 0.0f;
    a.best_cnt = 0;
    a.best_id = a.best_by = 0;
    a.cpsms = a.flags = 0;
    a.nia = a.nprobe = 0;
}

// Inclusive scan over the 64 lanes of a wavefront with DPP moves (no LDS
// round trips): shifts by 1, 2, 4, 8 within each 16-lane row (lanes shifted
// in from outside the row read 0), then row 0's last lane is added to row 1
// and row 2's to row 3 (row_bcast:15), and lane 31's to rows 2 and 3
// (row_bcast:31).
__device__ __forceinline__ uint32_t wave_scan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

// Reduction over the 64 lanes of a wavefront, complete in lane 63 (the
// DPP pattern of wave_scan with any associative op whose identity is 0).
template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce_last(uint32_t x, Op op) {
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false));
    return x;
}
__device__ __forceinline__ uint32_t wave_max_last(uint32_t x) {
    return wave_reduce_last(x, [](uint32_t u, uint32_t v) { return max(u, v); });
}

// Workgroup-wide exclusive scan where thread t holds tsum, the sum of its own
// contiguous group; returns t's prefix.  wsum: BT / 64 words; one barrier.
// After it each wavefront reads the wavefront sums once (a lane each) and
// scans them across its lanes, instead of every thread adding them up.
template <int BT>
__device__ __forceinline__ uint32_t block_prefix(lu32 *wsum, uint32_t tsum) {
    static_assert(BT / 64 <= 64, "one lane per wavefront sum");
    constexpr int NW = BT / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t x = wave_scan(tsum);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    const uint32_t sx = wave_scan(lane < NW ? (uint32_t)wsum[lane] : 0u);
    const uint32_t before = w ? (uint32_t)__builtin_amdgcn_readlane((int)sx, w - 1) : 0u;
    return x - tsum + before;
}

// Workgroup-wide exclusive scan of two values at once (thread t holds the
// sums of its own contiguous group); returns t's prefixes, totals in *tot.
// wsum: 2 * BT / 64 words; one barrier.
template <int BT>
__device__ __forceinline__ uint2 block_prefix2(lu32 *wsum, uint32_t a0, uint32_t a1, uint2 *tot) {
    static_assert(BT / 64 <= 64, "one lane per wavefront sum");
    constexpr int NW = BT / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t x = wave_scan(a0), y = wave_scan(a1);
    if (lane == 63) {
        wsum[w] = x;
        wsum[NW + w] = y;
    }
    __syncthreads();
    const uint32_t sx = wave_scan(lane < NW ? (uint32_t)wsum[lane] : 0u);
    const uint32_t sy = wave_scan(lane < NW ? (uint32_t)wsum[NW + lane] : 0u);
    const uint32_t bx = w ? (uint32_t)__builtin_amdgcn_readlane((int)sx, w - 1) : 0u;
    const uint32_t by = w ? (uint32_t)__builtin_amdgcn_readlane((int)sy, w - 1) : 0u;
    *tot = make_uint2((uint32_t)__builtin_amdgcn_readlane((int)sx, NW - 1),
                      (uint32_t)__builtin_amdgcn_readlane((int)sy, NW - 1));
    return make_uint2(x - a0 + bx, y - a1 + by);
}

// The run-end bitmap (kSegWords 64-bit words) for segment [s0, s0 +
// kSegEntries) of a run table whose flat offsets are cofs[0..nc] (run r ends
// at cofs[r + 1] - 1).
template <int BT>
__device__ __forceinline__ void build_seg(lu32 *seg32, const lu32 *cofs, uint32_t nc, uint32_t s0) {
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < 2 * kSegWords; w += BT) seg32[w] = 0;
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < nc; r += BT) {
        const uint32_t f = cofs[r + 1] - 1 - s0;
        if (f < kSegEntries) lds_or(&seg32[f >> 5], 1u << (f & 31u));
    }
    __syncthreads();
}

// The run holding flat index x: the number of runs that end before x.
__device__ __forceinline__ uint32_t run_at(const lu32 *cofs, uint32_t nc, uint32_t x) {
    uint32_t lo = 0, hi = nc;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (cofs[m + 1] <= x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// Walk the flat entries [s0, se) of the segment in the bitmap: wavefront w
// takes a contiguous share of its 64-entry tiles, U tiles per step, lane
// j of a tile the entry F + j.  The entry's run is the number of runs that
// end before it: the count before the tile (C, carried from tile to tile)
// plus the tile word's bits below j -- t.C + mbcnt(t.m), at
// most nc.  Per step, load(t, x, ok, v, fs) fills v[] for the U tiles
// (ok[u]: x[u] is in the wavefront's share; fs: the step's first flat index)
// and body(t, v, x, ok, fs) consumes them, so that each phase's memory
// operations of the U tiles issue back to back.
struct Tile {
    unsigned long long m;   // run ends in the tile, uniform
    uint32_t C;             // runs ending before the tile, uniform
};
__device__ __forceinline__ uint32_t tile_below(const Tile &t) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(t.m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)t.m, 0u));
}
// arr[run of this lane's entry]: the tile's base run is uniform, so the lane
// part is a scaled add to a scalar base
__device__ __forceinline__ uint32_t tile_at(const lu32 *arr, const Tile &t) { return (arr + t.C)[tile_below(t)]; }
template <int BT, uint32_t U, typename Load, typename Body>
__device__ __forceinline__ void walk_tiles(const lu64 *seg64, const lu32 *cofs, uint32_t nc, uint32_t s0, uint32_t se,
                                           Load &&load, Body &&body) {
    constexpr uint32_t NW = BT / 64;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t T = (se - s0 + 63) >> 6;
    const uint32_t k0 = ufl(T * w / NW), k1 = ufl(T * (w + 1) / NW);
    if (k0 >= k1) return;
    const uint32_t xe = min(se, s0 + 64 * k1);   // end of this wavefront's share
    // runs ending before the share: in segment 0, the bitmap words before it
    // (popcounts summed across the wavefront), else a binary search
    uint32_t C;
    if (s0 == 0) {
        uint32_t cnt = 0;
        for (uint32_t i = lane; i < k0; i += 64) cnt += (uint32_t)__popcll(seg64[i]);
        C = ufl((uint32_t)__builtin_amdgcn_readlane((int)wave_scan(cnt), 63));
    } else {
        C = ufl(run_at(cofs, nc, s0 + 64 * k0));
    }
    // a step whose U tiles lie wholly in the share (all but the last)
    // runs with ok known true
    auto step = [&](uint32_t k, auto whole_tag) {
        constexpr bool WHOLE = decltype(whole_tag)::value;
        Tile t[U];
        uint32_t x[U], v[U];
        bool ok[U];
        unsigned long long mv[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) mv[u] = seg64[WHOLE ? k + u : min(k + u, kSegWords - 1)];
        const uint32_t x0 = s0 + 64 * k + lane;
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            unsigned long long m = ((unsigned long long)ufl((uint32_t)(mv[u] >> 32)) << 32) | ufl((uint32_t)mv[u]);
            if (!WHOLE) m = k + u < k1 ? m : 0ull;
            t[u].m = m;
            t[u].C = C;
            C += (uint32_t)__popcll(m);
            x[u] = x0 + 64 * u;
            ok[u] = WHOLE ? true : x[u] < xe;
        }
        load(t, x, ok, v, s0 + 64 * k);
        body(t, v, x, ok, s0 + 64 * k);
    };
    uint32_t k = k0;
    for (; k + U <= k1 && s0 + 64 * (k + U) <= xe; k += U) step(k, std::true_type{});
    for (; k < k1; k += U) step(k, std::false_type{});
}

__device__ __forceinline__ LenDesc load_len(const ChunkArgs &a, int ixx) { return a.q.lens[ixx]; }

// the descriptor of length ixx, ixx uniform
__device__ __forceinline__ LenDesc load_len_u(const ChunkArgs &a, uint32_t ixx) {
    LenDesc L = a.q.lens[ufl(ixx)];
    L.mass = ufl(L.mass);
    L.iA = ufl(L.iA);
    L.bA = ufl(L.bA);
    L.mdir = ufl(L.mdir);
    L.n = ufl(L.n);
    L.speclen = ufl(L.speclen);
    L.half = ufl(L.half);
    L.peplen = ufl(L.peplen);
    L.div_m = ufl(L.div_m);
    L.div_s = ufl(L.div_s);
    L.nch = ufl(L.nch);
    L.cbase = ufl(L.cbase);
    L.bA2 = ufl(L.bA2);
    L.iAc = ufl(L.iAc);
    return L;
}

// A workgroup's accumulators for spectrum s merged into the global SpecAcc
// and the spectrum's histogram (hist: the low kHistLds bins, cleared here).
// s_red: [0] max bits, [1] count at max, [2] cpsms, [3] flags; s_red64:
// n_ia, n_probe.  Wavefront reductions first: one LDS atomic per wavefront and
// field (same-address atomics serialise).
template <int BT>
__device__ __forceinline__ void flush_spectrum(const ChunkArgs &a, lu32 *hist, uint32_t s, Acc &acc, lu32 *s_red,
                                               lu64 *s_red64) {
    if (threadIdx.x == 0) {
        s_red[0] = 0;
        s_red[1] = 0;
        s_red[2] = 0;
        s_red[3] = 0;
        s_red64[0] = 0;
        s_red64[1] = 0;
    }
    __syncthreads();
    {
        auto add = [](uint32_t u, uint32_t v) { return u + v; };
        const uint32_t hm = wave_max_last(acc.best_cnt ? __float_as_uint(acc.best_hs) : 0u);
        const uint32_t cp = wave_reduce_last(acc.cpsms, add);
        const uint32_t fl = wave_reduce_last(acc.flags, [](uint32_t u, uint32_t v) { return u | v; });
        const uint32_t ni = wave_reduce_last(acc.nia, add);
        const uint32_t np = wave_reduce_last(acc.nprobe, add);
        if ((threadIdx.x & 63u) == 63u) {
            if (hm) lds_max(&s_red[0], hm);
            if (cp) lds_add(&s_red[2], cp);
            if (fl) lds_or(&s_red[3], fl);
            if (ni) lds_add64(&s_red64[0], (unsigned long long)ni);
            if (np) lds_add64(&s_red64[1], (unsigned long long)np);
        }
    }
    __syncthreads();
    SpecAcc *g = a.acc + s;
    const uint32_t mb = s_red[0];
    if (acc.best_cnt && __float_as_uint(acc.best_hs) == mb) {
        lds_add(&s_red[1], acc.best_cnt);
        glb_max64(&g->top_a, ((unsigned long long)mb << 32) | acc.best_id);
        glb_max64(&g->top_b, ((unsigned long long)mb << 32) | acc.best_by);
    }
    // histogram: non-zero bins to global, then clear
    uint32_t *hg = a.q.hist + (size_t)s * a.q.hbins;
    for (uint32_t h = threadIdx.x; h < min(a.q.hbins, kHistLds); h += BT) {
        const uint32_t v = hist[h];
        if (v) {
            glb_add(&hg[h], v);
            hist[h] = 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_red[2]) glb_add(&g->cpsms, (uint32_t)s_red[2]);
        if (s_red[3]) glb_or(&g->flags, s_red[3]);
        if (s_red64[0]) glb_add(&g->n_ia, (unsigned long long)s_red64[0]);
        if (s_red64[1]) glb_add(&g->n_probe, (unsigned long long)s_red64[1]);
        const uint32_t cnt = s_red[1];
        if (cnt) {
            gu64 *tn = (gu64 *)&g->top_n;
            const unsigned long long mine = ((unsigned long long)mb << 32) | cnt;
            unsigned long long old = *tn;
            for (;;) {
                const uint32_t oh = (uint32_t)(old >> 32);
                if (mb < oh) break;
                const unsigned long long want = (mb > oh) ? mine : old + cnt;
                if (__hip_atomic_compare_exchange_strong(tn, &old, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT))
                    break;
            }
        }
    }
    __syncthreads();
    acc_reset(acc);
}

// The peaks of spectrum s into LDS; returns their count, uniform.
template <int BT>
__device__ __forceinline__ uint32_t load_peaks(const ChunkArgs &a, lu32 *pk_moz, lu32 *pk_int, uint32_t s) {
    gcu32 *qidx = (gcu32 *)a.q.qidx, *moz = (gcu32 *)a.q.moz, *inten = (gcu32 *)a.q.inten;
    const uint32_t q0 = qidx[s];
    const uint32_t P = ufl(qidx[s + 1] - q0);
    for (uint32_t k = threadIdx.x; k < P; k += BT) {
        pk_moz[k] = moz[q0 + k];
        pk_int[k] = inten[q0 + k];
    }
    return P;
}

// ---- the batched kernels' work items (ChunkArgs) ----

// Spectrum s's windows into LDS (wlo / whi, nlen each) and, per length, its
// first work item within the spectrum (lfirst, nlen + 1 entries; lfirst[nlen]
// = the spectrum's item count).  A length's items are the chunks its window
// touches.  Readers must be past a barrier.
template <int BT>
__device__ __forceinline__ void load_windows(const ChunkArgs &a, uint32_t s, li32 *wlo, li32 *whi, lu32 *lfirst) {
    const int nlen = a.q.nlen;
    const int2 *win = a.win + (size_t)s * nlen;
    if (threadIdx.x < 64) {
        uint32_t carry = 0;
        for (int l0 = 0; l0 < nlen; l0 += 64) {
            const int l = l0 + (int)threadIdx.x;
            uint32_t cnt = 0;
            if (l < nlen) {
                const int2 w = win[l];
                wlo[l] = w.x;
                whi[l] = w.y;
                if (w.x <= w.y) cnt = (uint32_t)w.y / a.tc - (uint32_t)w.x / a.tc + 1;
            }
            const uint32_t inc = wave_scan(cnt);
            if (l < nlen) lfirst[l] = carry + inc - cnt;
            carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        }
        if (threadIdx.x == 0) lfirst[nlen] = carry;
    }
}

// The spectrum holding item `it`: item_spec gives the spectrum holding item
// 64 * (it / 64), and spectra are numbered in item order.  Uniform.
__device__ __forceinline__ uint32_t spec_of_item(const ChunkArgs &a, uint32_t it) {
    uint32_t s = ufl(((gcu32 *)a.item_spec)[it >> 6]);
    while (ufl(((gcu32 *)a.item_base)[s + 1]) <= it) s++;
    return s;
}

// Item k of the current spectrum (k < lfirst[nlen]): its length *l and chunk
// *c.  Lengths without items have lfirst[l + 1] == lfirst[l].  Uniform.
__device__ __forceinline__ void item_at(const ChunkArgs &a, const li32 *wlo, const lu32 *lfirst, uint32_t k,
                                        uint32_t *l, uint32_t *c) {
    // the lengths whose items end at or before k are a prefix; k is in the
    // first length after it (every wavefront counts them with ballots)
    const int nlen = a.q.nlen;
    uint32_t ll = 0;
    for (int l0 = 0; l0 < nlen; l0 += 64) {
        const int li = l0 + (int)(threadIdx.x & 63u);
        ll += (uint32_t)__popcll(__ballot(li < nlen && lfirst[li + 1] <= k));
    }
    ll = ufl(ll);
    *l = ll;
    *c = ufl((uint32_t)wlo[ll]) / a.tc + (k - ufl((uint32_t)lfirst[ll]));
}

// The persistent loop of the batched kernels: workgroups take runs of up to
// a.grab consecutive items from the queue (guided: about a quarter of a
// workgroup's share of what is left, so the workgroups finish together).
// begin(s) makes spectrum s current (after flush(previous) when the
// spectrum changes); item(s, l, c, imax_out) processes one item and records
// its best score in imax_out.  The grab's best scores collect in LDS (imax,
// kMaxGrab words) and go out together; s_item: one LDS word.
template <int BT, typename Begin, typename Item, typename Flush>
__device__ __forceinline__ void run_items(const ChunkArgs &a, lu32 *s_item, lu32 *imax, const li32 *wlo,
                                          const lu32 *lfirst, Begin &&begin, Item &&item, Flush &&flush) {
    const uint32_t total = ufl(((gcu32 *)a.item_base)[a.q.nspec]);
    const uint32_t gmax = max(1u, min(a.grab, kMaxGrab));
    uint32_t grab = max(1u, min(gmax, total / (4u * gridDim.x)));
    uint32_t cur = 0xFFFFFFFFu, cbase = 0, cend = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0)
            *s_item = __hip_atomic_fetch_add((gu32 *)a.queue, grab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t it0 = ufl((uint32_t)*s_item);
        if (it0 >= total) break;
        const uint32_t it1 = min(total, it0 + grab);
        grab = max(1u, min(gmax, (total - it1) / (4u * gridDim.x)));
        const bool rec = a.item_max != nullptr;
        if (rec && threadIdx.x < it1 - it0) imax[threadIdx.x] = 0;
        uint32_t s = (cur != 0xFFFFFFFFu && it0 >= cbase && it0 < cend) ? cur : spec_of_item(a, it0);
        for (uint32_t it = it0; it < it1; it++) {
            while (it >= cend || s != cur) {
                if (s == cur) s++;   // past the current spectrum: the next with items
                const uint32_t b0 = ufl(((gcu32 *)a.item_base)[s]), b1 = ufl(((gcu32 *)a.item_base)[s + 1]);
                if (it >= b1) {
                    s++;
                    continue;
                }
                if (cur != 0xFFFFFFFFu) flush(cur);
                cur = s;
                cbase = b0;
                cend = b1;
                begin(s);
                __syncthreads();
            }
            uint32_t l, c;
            item_at(a, wlo, lfirst, it - cbase, &l, &c);
            item(cur, l, c, rec ? imax + (it - it0) : nullptr);
        }
        __syncthreads();
        if (rec && threadIdx.x < it1 - it0 && it0 + threadIdx.x < a.item_cap)
            ((AS_GLB uint32_t *)a.item_max)[it0 + threadIdx.x] = imax[threadIdx.x];
    }
    if (cur != 0xFFFFFFFFu) flush(cur);
}

}  // namespace
}  // namespace hcx
