// k_lbe.hip -- the peptide database stage on the GPU (LBE + MODS): the mass
// filter of LBE_CountPeps (lbe_internal.cpp:369-467), the variant counts of
// MODS_ModCounter (mods.cpp:399-432), the rank share of LBE_CreatePartitions
// (lbe_internal.cpp:308-352), the entries of LBE_GeneratePeps
// (lbe_internal.cpp:204-229) and MODS_GenerateMods (mods.cpp:448-496), and
// LBE_Initialize's sort by mass (lbe_internal.cpp:184-187).
//
// Orders that decide which row a form gets (and so every iA id and the tie
// order of top hits) are reproduced exactly:
//  * a sequence's variants come in MODS_ModList's depth-first order
//    (mods.cpp:333-384: modify the current residue before skipping it), and
//    glibc 2.35's qsort with cmpvarEntries is a stable merge sort by the site
//    distance of slm_dsts.h:191-271 (first modified residue s1 to last s3:
//    larger s1 + 60 - s3 first), i.e. a stable sort by ascending s3 - s1: each
//    thread enumerates its sequence twice, counting forms per span, then
//    writing each form to its slot;
//  * std::sort by mass (libstdc++ introsort: __introsort_loop with
//    median-of-three pivots moved to the range's first slot, the unguarded
//    Hoare partition, a heap sort once 2 floor(log2 n) levels are used up,
//    and a final insertion sort) leaves equal masses in an order fixed by its
//    partitions.  Each partition is computed in closed form: with G the
//    ascending positions after the pivot whose key is not below it and S the
//    descending positions whose key is not above it, the scans swap G[k] with
//    S[k] exactly for k < K = max_x min(#G before x, #S from x on), and stop at
//    min(G[K], S[K-1]) (G[0] when K = 0).  Disjoint ranges are partitioned side
//    by side: ranges above kSmall elements by one workgroup each over HBM,
//    smaller ones in LDS by one wavefront each, level by level.  Ranges of at
//    most 16 end in the final insertion sort, which never moves a key past a
//    range boundary (every key of a range is <= every key of the next), so it
//    is a stable sort inside each such range.
#include "hcx_internal.h"
#include "k_mass.h"
#include "k_scan_common.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <vector>

namespace hcx {

// Parsed modconditions (MODS_Initialize, mods.cpp:64-103).
struct LbeMods {
    int32_t limit;                        // mods per form, at most
    uint32_t ntypes;
    uint32_t cond[HCX_MAX_MOD_TYPES];     // per-type caps (clamped to 63)
    int32_t type_of[26];                  // residue 'A' + i -> type, -1: none
    uint32_t mod_mass[HCX_MAX_MOD_TYPES]; // (uint)(Da * scale)
    uint32_t scale;
};

struct SortRange {
    uint32_t first, last, depth, pad;
};

namespace {

constexpr uint32_t kSmall = 4096;   // ranges of at most this many keys are sorted in LDS
constexpr uint32_t kLeaf = 16;      // _S_threshold of libstdc++'s introsort
constexpr int kLT = 1024;           // threads of the large-range kernel
constexpr int kLE = 4;              // keys per thread per chunk
constexpr uint32_t kLChunk = kLT * kLE;
constexpr int kST = 256;            // threads of the small-range kernel
constexpr int kSW = kST / 64;
constexpr uint32_t kSmallList = kSmall / (kLeaf + 1) + 2;   // ranges > 16 at one level, at most
constexpr uint32_t kNone = 0xFFFFFFFFu;

__constant__ unsigned long long c_comb[64][64];   // mods.cpp:137-160

// ---- filter, counts, entries ----

__global__ void k_lbe_filter(char *lines, uint32_t n, uint32_t L, float min_mass, float max_mass, uint32_t *flag,
                             uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == n) flag[n] = 0;
    if (i >= n) return;
    char *s = lines + (size_t)i * L;
    bool bad = false;
    for (uint32_t l = 0; l < L; l++) {
        char c = s[l];
        if (c >= 'a' && c <= 'z') {
            c = (char)(c - 32);
            s[l] = c;
        }
        bad |= !(c >= 'A' && c <= 'Z');
    }
    if (bad) {
        atomicOr(err, 1u);
        flag[i] = 0;
        return;
    }
    const float m = pep_mass_dev(s, L);
    flag[i] = (m >= min_mass && m <= max_mass) ? 1u : 0u;
}

__global__ void k_lbe_compact(const char *lines, const uint32_t *flag, const uint32_t *pos, uint64_t nchars,
                              uint32_t L, char *seqs) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nchars) return;
    const uint32_t i = (uint32_t)(g / L), l = (uint32_t)(g - (uint64_t)i * L);
    if (flag[i]) seqs[(size_t)pos[i] * L + l] = lines[g];
}

// count() of mods.cpp:290-315 in closed form: partition3 equals the sum, over
// per-type counts k_t <= cap_t with sum k_t <= limit, of prod C(N_t, k_t),
// N_t the residues of type t in the sequence (Vandermonde on partition2).
// The kept count n is read on the device (the filter's scan total), so the
// host does not wait between the filter and the counts; the grid covers every
// input line and entries n..nmax are cleared for the scan that follows.
__global__ void k_lbe_count(const char *seqs, const uint32_t *np, uint32_t nmax, uint32_t L, LbeMods m,
                            uint32_t *vc, unsigned long long *total, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *np;
    unsigned long long v = 0;
    if (i >= n && i <= nmax) vc[i] = 0;
    if (i < n) {
        const char *s = seqs + (size_t)i * L;
        uint32_t N[HCX_MAX_MOD_TYPES];
        for (uint32_t t = 0; t < m.ntypes; t++) N[t] = 0;
        for (uint32_t l = 0; l < L; l++) {
            const int32_t t = m.type_of[s[l] - 'A'];
            if (t >= 0) N[t]++;
        }
        const uint32_t lim = (uint32_t)min(m.limit, 63);
        unsigned long long ways[64];
        ways[0] = 1;
        uint32_t top = 0;   // largest total with ways != 0 so far
        for (uint32_t t = 0; t < m.ntypes; t++) {
            const uint32_t kmax = min(min(m.cond[t], N[t]), lim);
            const uint32_t ntop = min(top + kmax, lim);
            for (uint32_t s2 = top + 1; s2 <= ntop; s2++) ways[s2] = 0;
            for (int32_t s2 = (int32_t)ntop; s2 >= 0; s2--) {
                unsigned long long acc = 0;
                const uint32_t kk = min(kmax, (uint32_t)s2);
                for (uint32_t k = 0; k <= kk; k++)
                    if ((uint32_t)s2 - k <= top) acc += ways[s2 - k] * c_comb[N[t]][k];
                ways[s2] = acc;
            }
            top = ntop;
        }
        unsigned long long forms = 0;
        for (uint32_t s2 = 0; s2 <= top; s2++) forms += ways[s2];
        v = m.limit > 0 ? forms - 1 : 0;
        if (v > 0xFFFFFFFFull) {
            atomicOr(err, 2u);
            v = 0;
        }
        vc[i] = (uint32_t)v;
    }
    // one atomic per wavefront
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(total, v);
}

__device__ __forceinline__ uint32_t generate_index(uint32_t key, uint32_t lclpep, uint32_t nodes, uint32_t myid,
                                                   int32_t policy) {
    // dslim.cpp:766-780 DSLIM_GenerateIndex
    return policy == HCX_POLICY_CHUNK ? myid * lclpep + key : key * nodes + myid;
}

__global__ void k_lbe_unmod(const char *seqs, uint32_t L, uint32_t lclpep, uint32_t nodes, uint32_t myid,
                            int32_t policy, hcx_pep_entry *out, float *keys, uint32_t *vals, uint32_t vbase) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= lclpep) return;
    const uint32_t idd = generate_index(f, lclpep, nodes, myid, policy);
    hcx_pep_entry e;
    e.mass = pep_mass_dev(seqs + (size_t)idd * L, L);
    e.seq_id = idd;
    e.sites = 0;
    e.mod_num = 0;
    e.pad_ = 0;
    out[f] = e;
    if (keys) {
        keys[f] = e.mass;
        vals[f] = vbase + f;
    }
}

// MODS_ModList (mods.cpp:333-384) without recursion: the chosen sites are the
// stack (the last choice is the highest set bit); at each step the next
// residue at or after `next` that its type's remaining cap allows is taken and
// the form emitted, else the last choice is dropped and the scan resumes after
// it.  Per-type remaining caps are bytes of two 64-bit words.
template <class Emit>
__device__ __forceinline__ void mod_forms(const char *s, uint32_t L, const LbeMods &m, const int32_t *ty,
                                          Emit emit) {
    unsigned long long capw[2] = {0, 0};
    for (uint32_t t = 0; t < m.ntypes; t++) capw[t >> 3] |= (unsigned long long)m.cond[t] << ((t & 7) * 8);
    auto cap = [&](int32_t t) -> uint32_t { return (uint32_t)(capw[t >> 3] >> ((t & 7) * 8)) & 0xFFu; };
    auto cap_add = [&](int32_t t, long long d) {
        capw[t >> 3] = (unsigned long long)((long long)capw[t >> 3] + (d << ((t & 7) * 8)));
    };
    unsigned long long sites = 0;
    uint32_t modnum = 0;
    int32_t d = 0;
    uint32_t next = 0;
    while (true) {
        int32_t found = -1, ft = -1;
        if (d < m.limit) {
            for (uint32_t l = next; l < L; l++) {
                const int32_t t = ty[l];
                if (t >= 0 && cap(t) > 0) {
                    found = (int32_t)l;
                    ft = t;
                    break;
                }
            }
        }
        if (found >= 0) {
            cap_add(ft, -1);
            sites |= 1ull << found;
            modnum += (uint32_t)(ft + 1) << (4 * d);
            d++;
            next = (uint32_t)found + 1;
            emit(sites, modnum);
            continue;
        }
        if (d == 0) break;
        const int32_t lp = 63 - __clzll((long long)sites);
        const int32_t t = ty[lp];
        cap_add(t, 1);
        sites &= ~(1ull << lp);
        d--;
        modnum -= (uint32_t)(t + 1) << (4 * d);
        next = (uint32_t)lp + 1;
    }
}

constexpr int kModT = 128;   // threads of k_lbe_mods: a 64-span histogram column each in LDS

__global__ __launch_bounds__(kModT) void k_lbe_mods(const char *seqs, uint32_t n, uint32_t L, LbeMods m,
                                                    const uint32_t *var_off, uint32_t lclpep, uint32_t lclmod,
                                                    uint32_t nodes, uint32_t myid, int32_t policy,
                                                    hcx_pep_entry *out, float *keys, uint32_t *vals,
                                                    uint32_t vbase, uint32_t *err) {
    __shared__ uint32_t hist[64][kModT];
    const uint32_t t = threadIdx.x;
    const uint32_t i = blockIdx.x * kModT + t;
    if (i >= n) return;
    const uint32_t g0 = var_off[i];
    if (var_off[i + 1] == g0) return;
    const char *s = seqs + (size_t)i * L;
    int32_t ty[64];
    for (uint32_t l = 0; l < L; l++) ty[l] = m.type_of[s[l] - 'A'];
    for (int k = 0; k < 64; k++) hist[k][t] = 0;
    auto keep = [&](uint32_t g) {
        // LBE_ApplyPolicy (lbe_internal.cpp:75-103) for modified forms
        return policy == HCX_POLICY_CHUNK ? (g / lclmod) == myid : (g % nodes) == myid;
    };
    uint32_t g = g0;
    mod_forms(s, L, m, ty, [&](unsigned long long sites, uint32_t) {
        if (keep(g)) {
            const uint32_t span = (uint32_t)(63 - __clzll((long long)sites)) - (uint32_t)__ffsll((long long)sites) + 1;
            hist[span][t]++;
        }
        g++;
    });
    if (g != var_off[i + 1]) {
        // the enumeration and the closed-form count disagree: write nothing
        atomicOr(err, 4u);
        return;
    }
    uint32_t run = 0;
    for (int k = 0; k < 64; k++) {
        const uint32_t c = hist[k][t];
        hist[k][t] = run;
        run += c;
    }
    const uint32_t local = g0 / nodes + (myid < g0 % nodes ? 1u : 0u);
    const float pm = pep_mass_dev(s, L);
    g = g0;
    mod_forms(s, L, m, ty, [&](unsigned long long sites, uint32_t modnum) {
        if (keep(g)) {
            const uint32_t span = (uint32_t)(63 - __clzll((long long)sites)) - (uint32_t)__ffsll((long long)sites) + 1;
            const uint32_t slot = lclpep + local + hist[span][t]++;
            if (slot >= lclpep + lclmod) {
                atomicOr(err, 4u);
                return;
            }
            // UTILS_CalculateModMass (utils.cpp:337-362): the first mod only
            const uint32_t first = modnum & 0xFu;
            hcx_pep_entry e;
            e.mass = pm + (float)m.mod_mass[first - 1] / (float)m.scale;
            e.seq_id = i;
            e.sites = sites;
            e.mod_num = modnum;
            e.pad_ = 0;
            out[slot] = e;
            if (keys) {
                keys[slot] = e.mass;
                vals[slot] = vbase + slot;
            }
        }
        g++;
    });
}

__global__ void k_lbe_gather(const hcx_pep_entry *src, const float *keys, const uint32_t *vals, uint32_t n,
                             hcx_pep_entry *out, float *mass) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = src[vals[i]];
    mass[i] = keys[i];
}

__global__ void k_lbe_mass(const hcx_pep_entry *e, uint32_t n, float *mass) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mass[i] = e[i].mass;
}

// ---- the exact introsort ----

// std::__move_median_to_first(first, first + 1, mid, last - 1)
template <typename KP, typename VP>
__device__ __forceinline__ void median_to_first(KP K, VP V, uint32_t r, uint32_t a, uint32_t b, uint32_t c) {
    const float ka = K[a], kb = K[b], kc = K[c];
    uint32_t x;
    if (ka < kb) {
        if (kb < kc) x = b;
        else if (ka < kc) x = c;
        else x = a;
    } else if (ka < kc) {
        x = a;
    } else if (kb < kc) {
        x = c;
    } else {
        x = b;
    }
    const float tk = K[r];
    const uint32_t tv = V[r];
    K[r] = K[x];
    V[r] = V[x];
    K[x] = tk;
    V[x] = tv;
}

// std::__adjust_heap + std::__push_heap on [K, K + len)
template <typename KP, typename VP>
__device__ void adjust_heap(KP K, VP V, int64_t hole, int64_t len, float kv, uint32_t vv) {
    const int64_t top = hole;
    int64_t second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (K[second] < K[second - 1]) second--;
        K[hole] = K[second];
        V[hole] = V[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        K[hole] = K[second - 1];
        V[hole] = V[second - 1];
        hole = second - 1;
    }
    int64_t parent = (hole - 1) / 2;
    while (hole > top && K[parent] < kv) {
        K[hole] = K[parent];
        V[hole] = V[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    K[hole] = kv;
    V[hole] = vv;
}

// std::__partial_sort(first, last, last): __make_heap, then __sort_heap
template <typename KP, typename VP>
__device__ void heap_sort_range(KP K, VP V, int64_t len) {
    if (len < 2) return;
    for (int64_t parent = (len - 2) / 2;; parent--) {
        adjust_heap(K, V, parent, len, K[parent], V[parent]);
        if (parent == 0) break;
    }
    for (int64_t last = len; last > 1;) {
        last--;
        const float kv = K[last];
        const uint32_t vv = V[last];
        K[last] = K[0];
        V[last] = V[0];
        adjust_heap(K, V, 0, last, kv, vv);
    }
}

__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o));
    return x;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, o));
    return x;
}

// Children of a partitioned range: leaves stay for the insertion sort (the
// small kernel sorts them), ranges of at most kSmall keys go to the small
// list, larger ones to the next level or, with the depth used up, to the heap
// list.  Capacities are checked; an overflow sets cnt[3].
__device__ void push_child(uint32_t first, uint32_t last, uint32_t depth, SortRange *large, uint32_t cap_large,
                           SortRange *small, uint32_t cap_small, SortRange *heap, uint32_t cap_heap, uint32_t *cnt) {
    const uint32_t n = last - first;
    SortRange r{first, last, depth, 0};
    uint32_t *c;
    SortRange *dst;
    uint32_t cap;
    if (n <= kSmall) {
        c = &cnt[1];
        dst = small;
        cap = cap_small;
    } else if (depth == 0) {
        c = &cnt[2];
        dst = heap;
        cap = cap_heap;
    } else {
        c = &cnt[0];
        dst = large;
        cap = cap_large;
    }
    const uint32_t k = atomicAdd(c, 1u);
    if (k < cap) dst[k] = r;
    else atomicOr(&cnt[3], 1u);
}

// One partition of a range above kSmall keys by one workgroup over HBM (the
// closed form of the file header), then its children pushed.  posS: the
// range's S positions by rank, at posS[first + rank].
__global__ __launch_bounds__(kLT) void k_sort_large(float *K, uint32_t *V, uint32_t *posS, const SortRange *in,
                                                    SortRange *out_large, uint32_t cap_large, SortRange *small,
                                                    uint32_t cap_small, SortRange *heap, uint32_t cap_heap,
                                                    uint32_t *cnt) {
    __shared__ uint32_t wsum[2 * kLT / 64];
    __shared__ uint32_t sh[4];
    __shared__ float piv;
    const SortRange r = in[blockIdx.x];
    const uint32_t f = r.first, l = r.last;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        median_to_first(K, V, f, f + 1, f + (l - f) / 2, l - 1);
        piv = K[f];
        sh[0] = 0;
        sh[1] = kNone;
        sh[2] = kNone;
    }
    __syncthreads();
    const float P = piv;
    // 1. totals of G (key >= P) and S (key <= P) over [f + 1, l)
    uint32_t cg = 0, cs = 0;
    for (uint32_t p = f + 1 + t; p < l; p += kLT) {
        const float k = K[p];
        cg += !(k < P);
        cs += !(P < k);
    }
    uint2 tot;
    (void)block_prefix2<kLT>((lu32 *)wsum, cg, cs, &tot);
    const uint32_t Gt = tot.x, St = tot.y;
    __syncthreads();
    // 2. K = max over x of min(#G in [f + 1, x), #S in [x, l)), from the left
    //    until the counts cross
    uint32_t aG = 0, aS = 0, best = 0;
    for (uint32_t base = f + 1; base < l; base += kLChunk) {
        uint32_t gm = 0, sm = 0;
        const uint32_t p0 = base + t * kLE;
#pragma unroll
        for (int j = 0; j < kLE; j++) {
            const uint32_t p = p0 + j;
            if (p < l) {
                const float k = K[p];
                gm |= (uint32_t)!(k < P) << j;
                sm |= (uint32_t)!(P < k) << j;
            }
        }
        const uint2 pre = block_prefix2<kLT>((lu32 *)wsum, __popc(gm), __popc(sm), &tot);
        uint32_t eg = aG + pre.x, es = aS + pre.y;
#pragma unroll
        for (int j = 0; j < kLE; j++) {
            if (p0 + j < l) best = max(best, min(eg, St - es));
            eg += (gm >> j) & 1u;
            es += (sm >> j) & 1u;
        }
        aG += tot.x;
        aS += tot.y;
        __syncthreads();
        if (aG >= St - aS) break;
    }
    best = max(best, min(aG, St - aS));
    best = wave_max_u32(best);
    if ((t & 63) == 0) atomicMax(&sh[0], best);
    __syncthreads();
    const uint32_t Kp = sh[0];
    // 3. S positions of rank < K from the right; S[K-1]
    if (Kp > 0) {
        uint32_t bS = 0;
        for (uint32_t end = l; end > f + 1;) {
            const uint32_t start = end - min(end - (f + 1), kLChunk);
            uint32_t sm = 0;
#pragma unroll
            for (int j = 0; j < kLE; j++) {
                const uint32_t q = t * kLE + j;
                if (q < end - start) sm |= (uint32_t)!(P < K[end - 1 - q]) << j;
            }
            const uint2 pre = block_prefix2<kLT>((lu32 *)wsum, 0u, __popc(sm), &tot);
            uint32_t rk = bS + pre.y;
#pragma unroll
            for (int j = 0; j < kLE; j++) {
                if ((sm >> j) & 1u) {
                    const uint32_t p = end - 1 - (t * kLE + j);
                    if (rk < Kp) posS[f + rk] = p;
                    if (rk == Kp - 1) sh[2] = p;
                    rk++;
                }
            }
            bS += tot.y;
            end = start;
            __syncthreads();
            if (bS >= Kp) break;
        }
    }
    __syncthreads();
    const uint32_t sK1 = sh[2];
    // 4. G positions of rank < K from the left, swapped with S of the same
    //    rank; G[K] if it lies before S[K-1]
    const uint32_t bound = Kp > 0 ? sK1 : l;
    aG = 0;
    for (uint32_t base = f + 1; base < bound; base += kLChunk) {
        uint32_t gm = 0;
        const uint32_t p0 = base + t * kLE;
#pragma unroll
        for (int j = 0; j < kLE; j++) {
            const uint32_t p = p0 + j;
            if (p < bound) gm |= (uint32_t)!(K[p] < P) << j;
        }
        const uint2 pre = block_prefix2<kLT>((lu32 *)wsum, __popc(gm), 0u, &tot);
        uint32_t rk = aG + pre.x;
#pragma unroll
        for (int j = 0; j < kLE; j++) {
            if ((gm >> j) & 1u) {
                const uint32_t p = p0 + j;
                if (rk < Kp) {
                    const uint32_t q = posS[f + rk];
                    const float kp = K[p], kq = K[q];
                    const uint32_t vp = V[p], vq = V[q];
                    K[p] = kq;
                    V[p] = vq;
                    K[q] = kp;
                    V[q] = vp;
                } else if (rk == Kp) {
                    sh[1] = p;
                }
                rk++;
            }
        }
        aG += tot.x;
        __syncthreads();
        if (aG > Kp) break;
    }
    __syncthreads();
    if (t == 0) {
        const uint32_t gK = sh[1];
        const uint32_t cut = Kp == 0 ? gK : min(gK, sK1);
        if (cut <= f || cut >= l) {
            // a cut outside the range cannot come from a partition of finite keys
            atomicOr(&cnt[3], 2u);
            return;
        }
        const uint32_t d = r.depth - 1;
        push_child(f, cut, d, out_large, cap_large, small, cap_small, heap, cap_heap, cnt);
        push_child(cut, l, d, out_large, cap_large, small, cap_small, heap, cap_heap, cnt);
    }
}

// A range above kSmall keys whose depth is used up: libstdc++'s heap sort by
// one lane over HBM (reached only by inputs that defeat median-of-three).
__global__ void k_sort_heap(float *K, uint32_t *V, const SortRange *list) {
    if (threadIdx.x != 0) return;
    const SortRange r = list[blockIdx.x];
    heap_sort_range(K + r.first, V + r.first, (int64_t)(r.last - r.first));
}

// The closed-form partition of [f, l) in LDS by one wavefront; returns the cut.
// pos: room for (l - f) / 2 positions.
__device__ uint32_t wave_partition(float *ks, uint32_t *vs, uint16_t *pos, uint32_t f, uint32_t l) {
    const uint32_t lane = threadIdx.x & 63;
    if (lane == 0) median_to_first(ks, vs, f, f + 1, f + (l - f) / 2, l - 1);
    __syncwarp();
    const float P = ks[f];
    uint32_t Gt = 0, St = 0;
    for (uint32_t b = f + 1; b < l; b += 64) {
        const uint32_t p = b + lane;
        const bool in = p < l;
        const float k = in ? ks[p] : 0.0f;
        Gt += __popcll(__ballot(in && !(k < P)));
        St += __popcll(__ballot(in && !(P < k)));
    }
    uint32_t aG = 0, aS = 0, best = 0;
    for (uint32_t b = f + 1; b < l; b += 64) {
        const uint32_t p = b + lane;
        const bool in = p < l;
        const float k = in ? ks[p] : 0.0f;
        const unsigned long long bg = __ballot(in && !(k < P)), bs = __ballot(in && !(P < k));
        const uint32_t eg = aG + lanes_below(bg), es = aS + lanes_below(bs);
        if (in) best = max(best, min(eg, St - es));
        aG += __popcll(bg);
        aS += __popcll(bs);
        if (aG >= St - aS) break;
    }
    best = max(best, min(aG, St - aS));
    const uint32_t Kp = wave_max_u32(best);
    uint32_t sK1 = kNone;
    if (Kp > 0) {
        uint32_t bS = 0;
        for (int32_t e = (int32_t)l; e > (int32_t)f + 1; e -= 64) {
            const int32_t p = e - 1 - (int32_t)lane;
            const bool in = p >= (int32_t)f + 1;
            const bool s = in && !(P < ks[in ? p : f]);
            const unsigned long long bs = __ballot(s);
            const uint32_t rk = bS + lanes_below(bs);
            if (s && rk < Kp) pos[rk] = (uint16_t)p;
            if (s && rk == Kp - 1) sK1 = (uint32_t)p;
            bS += __popcll(bs);
            if (bS >= Kp) break;
        }
        sK1 = wave_min_u32(sK1);
    }
    __syncwarp();
    const uint32_t bound = Kp > 0 ? sK1 : l;
    uint32_t gK = kNone;
    aG = 0;
    for (uint32_t b = f + 1; b < bound; b += 64) {
        const uint32_t p = b + lane;
        const bool in = p < bound;
        const float k = in ? ks[p] : 0.0f;
        const bool g = in && !(k < P);
        const unsigned long long bg = __ballot(g);
        const uint32_t rk = aG + lanes_below(bg);
        if (g && rk < Kp) {
            const uint32_t q = pos[rk];
            const float kq = ks[q];
            const uint32_t vp = vs[p], vq = vs[q];
            ks[p] = kq;
            vs[p] = vq;
            ks[q] = k;
            vs[q] = vp;
        }
        if (g && rk == Kp) gK = p;
        aG += __popcll(bg);
        if (aG > Kp) break;
    }
    gK = wave_min_u32(gK);
    __syncwarp();
    return Kp == 0 ? gK : min(gK, sK1);
}

// Stable sort of at most 16 keys at [f, l) by lanes 0..15 of a wavefront.
__device__ void wave_leaf_sort(float *ks, uint32_t *vs, uint32_t f, uint32_t l) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t m = l - f;
    const bool in = lane < m;
    const float k = in ? ks[f + lane] : 0.0f;
    const uint32_t v = in ? vs[f + lane] : 0u;
    uint32_t rank = 0;
    for (uint32_t j = 0; j < m; j++) {
        const float kj = __shfl(k, (int)j);
        rank += (kj < k) || (kj == k && j < lane);
    }
    __syncwarp();
    if (in) {
        ks[f + rank] = k;
        vs[f + rank] = v;
    }
    __syncwarp();
}

__device__ __forceinline__ uint32_t pack_range(uint32_t f, uint32_t l, uint32_t d) {
    return f | (l << 13) | (d << 26);
}

// One range of at most kSmall keys per workgroup, sorted in LDS to the end:
// partitions level by level, one wavefront per pending range; leaves stably
// sorted, ranges whose depth is used up heap-sorted.
__global__ __launch_bounds__(kST) void k_sort_small(float *K, uint32_t *V, const SortRange *list, uint32_t *cnt) {
    __shared__ float ks[kSmall];
    __shared__ uint32_t vs[kSmall];
    __shared__ uint16_t pos[kSW][kSmall / 2];
    __shared__ uint32_t lst[2][kSmallList];
    __shared__ uint32_t lcnt[2];
    const SortRange r = list[blockIdx.x];
    const uint32_t n = r.last - r.first;
    const uint32_t t = threadIdx.x, w = t >> 6;
    if (n > kSmall) {
        if (t == 0) atomicOr(&cnt[3], 2u);
        return;
    }
    for (uint32_t i = t; i < n; i += kST) {
        ks[i] = K[r.first + i];
        vs[i] = V[r.first + i];
    }
    if (t == 0) {
        lcnt[0] = 0;
        lcnt[1] = 0;
    }
    __syncthreads();
    if (n <= kLeaf) {
        if (w == 0) wave_leaf_sort(ks, vs, 0, n);
    } else if (r.depth == 0) {
        if (t == 0) heap_sort_range(ks, vs, (int64_t)n);
    } else if (t == 0) {
        lst[0][0] = pack_range(0, n, r.depth);
        lcnt[0] = 1;
    }
    __syncthreads();
    uint32_t cur = 0;
    while (true) {
        const uint32_t nc = min(lcnt[cur], kSmallList);
        if (nc == 0) break;
        __syncthreads();
        if (t == 0) lcnt[cur ^ 1] = 0;
        __syncthreads();
        for (uint32_t ri = w; ri < nc; ri += kSW) {
            const uint32_t pr = lst[cur][ri];
            const uint32_t f = pr & 0x1FFFu, l = (pr >> 13) & 0x1FFFu, d = pr >> 26;
            const uint32_t cut = wave_partition(ks, vs, pos[w], f, l);
            if (cut <= f || cut >= l) {
                if ((t & 63) == 0) atomicOr(&cnt[3], 2u);
                continue;
            }
            const uint32_t lo[2] = {f, cut}, hi[2] = {cut, l};
            for (int c = 0; c < 2; c++) {
                const uint32_t m = hi[c] - lo[c];
                if (m <= kLeaf) {
                    wave_leaf_sort(ks, vs, lo[c], hi[c]);
                } else if (d - 1 == 0) {
                    if ((t & 63) == 0) heap_sort_range(ks + lo[c], vs + lo[c], (int64_t)m);
                    __syncwarp();
                } else if ((t & 63) == 0) {
                    const uint32_t k = atomicAdd(&lcnt[cur ^ 1], 1u);
                    if (k < kSmallList) lst[cur ^ 1][k] = pack_range(lo[c], hi[c], d - 1);
                    else atomicOr(&cnt[3], 2u);
                }
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    __syncthreads();
    for (uint32_t i = t; i < n; i += kST) {
        K[r.first + i] = ks[i];
        V[r.first + i] = vs[i];
    }
}

uint32_t floor_log2(uint32_t n) { return 31u - (uint32_t)__builtin_clz(n); }

}  // namespace

// Sorts the segments [seg[i].x, seg[i].y) of (K, V) as std::sort would by key
// (each segment on its own).  Ends with the stream synchronised.
int sort_mass_segments(hcx_ctx *c, float *K, uint32_t *V, uint64_t total, const std::vector<uint2> &seg) {
    if (total > 0xFFFFFFF0ull) return set_error_code(HCX_ERR_INVLD_SIZE, "sort: more than 2^32 keys");
    hipStream_t st = c->stream;
    std::vector<SortRange> large, small;
    for (const uint2 &s : seg) {
        const uint32_t n = s.y - s.x;
        if (n < 2) continue;
        const SortRange r{s.x, s.y, 2 * floor_log2(n), 0};
        (n <= kSmall ? small : large).push_back(r);
    }
    const uint32_t nseg = (uint32_t)seg.size() + 1;
    const uint32_t cap_large = (uint32_t)(total / (kSmall + 1)) + nseg + 1;
    const uint64_t cap_small64 = 2ull * (62ull * (total / kSmall + 1) + nseg) + small.size() + 64;
    const uint32_t cap_small = (uint32_t)std::min<uint64_t>(cap_small64, total + nseg + 64);
    const uint32_t cap_heap = cap_large;
    DeviceBuf pos, la, lb, ls, lh, cnt;
    if (pos.ensure(4 * total + 64) || la.ensure(sizeof(SortRange) * cap_large) ||
        lb.ensure(sizeof(SortRange) * cap_large) || ls.ensure(sizeof(SortRange) * cap_small) ||
        lh.ensure(sizeof(SortRange) * cap_heap) || cnt.ensure(64))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "sort: device allocation failed");
    uint32_t h[4] = {0, (uint32_t)small.size(), 0, 0};
    if (!large.empty())
        HCX_HIP(hipMemcpyAsync(la.p, large.data(), sizeof(SortRange) * large.size(), hipMemcpyHostToDevice, st));
    if (!small.empty())
        HCX_HIP(hipMemcpyAsync(ls.p, small.data(), sizeof(SortRange) * small.size(), hipMemcpyHostToDevice, st));
    HCX_HIP(hipMemcpyAsync(cnt.p, h, 16, hipMemcpyHostToDevice, st));
    uint32_t nl = (uint32_t)large.size();
    SortRange *in = (SortRange *)la.p, *outl = (SortRange *)lb.p;
    while (nl > 0) {
        HCX_HIP(hipMemsetAsync(cnt.p, 0, 4, st));
        k_sort_large<<<nl, kLT, 0, st>>>(K, V, (uint32_t *)pos.p, in, outl, cap_large, (SortRange *)ls.p, cap_small,
                                         (SortRange *)lh.p, cap_heap, (uint32_t *)cnt.p);
        HCX_HIP(hipGetLastError());
        HCX_HIP(hipMemcpyAsync(h, cnt.p, 16, hipMemcpyDeviceToHost, st));
        HCX_HIP(hipStreamSynchronize(st));
        if (h[3]) return set_error_code(HCX_ERR_INVLD_SORT, "sort: internal error (range list or cut)");
        nl = h[0];
        std::swap(in, outl);
    }
    HCX_HIP(hipMemcpyAsync(h, cnt.p, 16, hipMemcpyDeviceToHost, st));
    HCX_HIP(hipStreamSynchronize(st));
    if (h[3] || h[1] > cap_small || h[2] > cap_heap) return set_error_code(HCX_ERR_INVLD_SIZE, "sort: range list overflow");
    if (h[2]) k_sort_heap<<<h[2], 64, 0, st>>>(K, V, (const SortRange *)lh.p);
    if (h[1]) k_sort_small<<<h[1], kST, 0, st>>>(K, V, (const SortRange *)ls.p, (uint32_t *)cnt.p);
    HCX_HIP(hipGetLastError());
    HCX_HIP(hipMemcpyAsync(h, cnt.p, 16, hipMemcpyDeviceToHost, st));
    HCX_HIP(hipStreamSynchronize(st));
    if (h[3]) return set_error_code(HCX_ERR_INVLD_SORT, "sort: internal error (range list or cut)");
    return HCX_SUCCESS;
}

}  // namespace hcx

// ---------------------------------------------------------------- C ABI ----

struct hcx_pepdb {
    hcx_ctx *ctx = nullptr;
    hcx_lbe_cfg cfg{};
    hcx::LbeMods mods{};
    uint32_t peplen = 0, nlines = 0, nseqs = 0, mod_count = 0, lclpep = 0, lclmod = 0;
    bool generated = false;
    float upload_ms = 0.0f, count_ms = 0.0f, generate_ms = 0.0f;
    hcx::DeviceBuf seqs, var_off, entries, mass;
};

using hcx::set_error;
using hcx::set_error_code;

namespace {

// MODS_Initialize (mods.cpp:64-103) with the checks the reference leaves to
// its input: residues A-Z, each in one condition only (the count and the
// enumeration would otherwise disagree), caps >= 0, at most 8 mods per form
// (modNum holds 8 nibbles), types within the context's mod table.
int parse_mods(const char *text, const hcx_params &p, hcx::LbeMods *m) {
    std::memset(m, 0, sizeof *m);
    for (int i = 0; i < 26; i++) m->type_of[i] = -1;
    std::memcpy(m->mod_mass, p.mod_mass, sizeof m->mod_mass);
    m->scale = p.scale;
    std::stringstream ss(text ? text : "0 M 0");
    std::vector<std::string> tok;
    std::string t;
    while (ss >> t) tok.push_back(t);
    if (tok.empty()) return set_error_code(HCX_ERR_INVLD_MOD, "modconditions: empty");
    char *end = nullptr;
    const long lim = std::strtol(tok[0].c_str(), &end, 10);
    if (!end || *end || lim < 0 || lim > 8)
        return set_error_code(HCX_ERR_INVLD_MOD, "modconditions: the mod limit must be 0..8");
    m->limit = (int32_t)lim;
    // pairs of (residues, count) after the limit; like MODS_Initialize
    // (mods.cpp:89-100), a trailing residue token without a count is ignored
    const size_t nt = (tok.size() - 1) / 2;
    if (nt > HCX_MAX_MOD_TYPES) return set_error_code(HCX_ERR_INVLD_MOD, "modconditions: too many mod types");
    m->ntypes = (uint32_t)nt;
    for (size_t i = 0; i < nt; i++) {
        for (char ch : tok[2 * i + 1]) {
            if (ch < 'A' || ch > 'Z') return set_error_code(HCX_ERR_INVLD_MOD, "modconditions: residues must be A-Z");
            if (m->type_of[ch - 'A'] != -1)
                return set_error_code(HCX_ERR_INVLD_MOD, "modconditions: a residue named twice");
            m->type_of[ch - 'A'] = (int32_t)i;
        }
        const long cap = std::strtol(tok[2 * i + 2].c_str(), &end, 10);
        if (!end || *end || cap < 0) return set_error_code(HCX_ERR_INVLD_MOD, "modconditions: bad per-type cap");
        m->cond[i] = (uint32_t)std::min<long>(cap, 63);
        if (m->limit > 0 && i >= p.nmods)
            return set_error_code(HCX_ERR_INVLD_MOD, "modconditions: more mod types than the context's mod table");
    }
    return HCX_SUCCESS;
}

// The binomial table into c_comb of the current device, once per device
// (callers on several threads may race to it: the flag is taken under a lock,
// the table is built on the caller's stack).
int comb_table_upload(int device) {
    static std::mutex mu;
    static bool done[64] = {false};
    if (device < 0 || device >= 64) return set_error("comb table: device id");
    std::lock_guard<std::mutex> g(mu);
    if (done[device]) return HCX_SUCCESS;
    // mods.cpp:137-160 MODS_GenCombinations
    std::vector<unsigned long long> tab(64 * 64, 0ull);
    auto C = reinterpret_cast<unsigned long long(*)[64]>(tab.data());
    C[0][0] = 1;
    C[1][0] = 1;
    C[1][1] = 1;
    for (int i = 1; i < 64; i++) {
        C[i][0] = 1;
        for (int j = 1; j <= (i + 1) / 2; j++) C[i][j] = C[i - 1][j - 1] + C[i - 1][j];
        for (int j = i / 2; j < i; j++) C[i][j] = C[i][i - j];
        C[i][i] = 1;
    }
    HCX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(hcx::c_comb), tab.data(), sizeof(unsigned long long) * 64 * 64));
    done[device] = true;
    return HCX_SUCCESS;
}

}  // namespace

extern "C" int hcx_pepdb_create(hcx_ctx *c, const hcx_lbe_cfg *cfg, uint32_t peplen, const char *lines,
                                uint32_t nlines, hcx_pepdb **out) {
    if (!c || !cfg || !out || (nlines && !lines)) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_pepdb_create: null");
    if (peplen < 2 || peplen > 60) return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_pepdb_create: bad peptide length");
    if (cfg->nodes == 0 || cfg->myid >= cfg->nodes)
        return set_error_code(HCX_ERR_INVLD_NODE_RANK, "hcx_pepdb_create: bad rank");
    if (cfg->policy != HCX_POLICY_CHUNK && cfg->policy != HCX_POLICY_CYCLIC)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_pepdb_create: unknown policy");
    if (cfg->policy == HCX_POLICY_CHUNK && cfg->nodes > 1)
        return set_error_code(HCX_ERR_INVLD_PARAM,
                              "hcx_pepdb_create: the chunk policy over several ranks places forms inconsistently");
    hcx_pepdb *db = new (std::nothrow) hcx_pepdb;
    if (!db) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_pepdb_create: out of memory");
    db->ctx = c;
    db->cfg = *cfg;
    db->cfg.modconditions = nullptr;
    db->peplen = peplen;
    db->nlines = nlines;
    auto fail = [&](int rc) {
        delete db;
        return rc;
    };
    if (int rc = parse_mods(cfg->modconditions, c->params, &db->mods)) return fail(rc);
    if (hipSetDevice(c->device) != hipSuccess) return fail(set_error("hipSetDevice"));
    if (int rc = comb_table_upload(c->device)) return fail(rc);
    hipStream_t st = c->stream;
    const uint64_t nchars = (uint64_t)nlines * peplen;
    hcx::DeviceBuf dl, flag, fpos, scan, err;
    // the kept sequences and their variant offsets are sized for every line,
    // so nothing is allocated between the upload and the counts
    if (dl.ensure(nchars + 16) || flag.ensure(4ull * nlines + 8) || fpos.ensure(4ull * nlines + 8) ||
        scan.ensure(4 * hcx::scan_tmp_words((uint64_t)nlines + 1) + 64) || err.ensure(64) ||
        db->seqs.ensure(nchars + 16) || db->var_off.ensure(4ull * nlines + 8))
        return fail(set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_pepdb_create: device allocation failed"));
    if (hipEventRecord(c->ev2, st) != hipSuccess) return fail(set_error("event"));
    if ((nchars && hipMemcpyAsync(dl.p, lines, nchars, hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemsetAsync(err.p, 0, 64, st) != hipSuccess || hipEventRecord(c->ev3, st) != hipSuccess)
        return fail(set_error("hcx_pepdb_create: upload failed"));
    hcx::k_lbe_filter<<<(nlines + 256) / 256, 256, 0, st>>>((char *)dl.p, nlines, peplen, (float)c->params.min_mass,
                                                            (float)c->params.max_mass, (uint32_t *)flag.p,
                                                            (uint32_t *)err.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error("hcx_pepdb_create: filter launch failed"));
    if (int rc = hcx::scan_exclusive_u32(st, (const uint32_t *)flag.p, (uint32_t *)fpos.p, (uint64_t)nlines + 1,
                                         (uint32_t *)scan.p))
        return fail(rc);
    if (nchars)
        hcx::k_lbe_compact<<<(unsigned)((nchars + 255) / 256), 256, 0, st>>>(
            (const char *)dl.p, (const uint32_t *)flag.p, (const uint32_t *)fpos.p, nchars, peplen, (char *)db->seqs.p);
    hcx::k_lbe_count<<<(nlines + 256) / 256, 256, 0, st>>>((const char *)db->seqs.p, (const uint32_t *)fpos.p + nlines,
                                                           nlines, peplen, db->mods, (uint32_t *)flag.p,
                                                           (unsigned long long *)((char *)err.p + 8), (uint32_t *)err.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error("hcx_pepdb_create: count launch failed"));
    if (int rc = hcx::scan_exclusive_u32(st, (const uint32_t *)flag.p, (uint32_t *)db->var_off.p,
                                         (uint64_t)nlines + 1, (uint32_t *)scan.p))
        return fail(rc);
    if (hipEventRecord(c->ev1, st) != hipSuccess) return fail(set_error("event"));
    uint32_t he[4] = {0, 0, 0, 0}, n = 0;
    if (hipMemcpyAsync(he, err.p, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&n, (const uint32_t *)fpos.p + nlines, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st))
        return fail(set_error("hcx_pepdb_create: count failed"));
    if (he[0] & 1u)
        return fail(set_error_code(HCX_ERR_INVLD_PARAM, "hcx_pepdb_create: sequence characters must be letters"));
    db->nseqs = n;
    const uint64_t mods = (uint64_t)he[2] | ((uint64_t)he[3] << 32);
    if (he[0] & 2u || mods > 0xFFFFFFFFull)
        return fail(set_error_code(HCX_ERR_INVLD_SIZE, "hcx_pepdb_create: more than 2^32 variants"));
    db->mod_count = (uint32_t)mods;
    // LBE_CreatePartitions (lbe_internal.cpp:308-352)
    const uint32_t p = cfg->nodes, id = cfg->myid;
    db->lclpep = n / p + ((n % p) > id ? 1u : 0u);
    db->lclmod = db->mod_count / p + ((db->mod_count % p) > id ? 1u : 0u);
    if ((uint64_t)db->lclpep + db->lclmod > 0xFFFFFFF0ull)
        return fail(set_error_code(HCX_ERR_INVLD_SIZE, "hcx_pepdb_create: more than 2^32 forms"));
    (void)hipEventElapsedTime(&db->upload_ms, c->ev2, c->ev3);
    (void)hipEventElapsedTime(&db->count_ms, c->ev3, c->ev1);
    *out = db;
    return HCX_SUCCESS;
}

extern "C" int hcx_pepdb_counts(const hcx_pepdb *db, uint32_t *counts) {
    if (!db || !counts) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_pepdb_counts: null");
    counts[0] = db->nseqs;
    counts[1] = db->mod_count;
    counts[2] = db->nseqs + db->mod_count;
    counts[3] = db->lclpep;
    counts[4] = db->lclmod;
    counts[5] = db->lclpep + db->lclmod;
    return HCX_SUCCESS;
}

extern "C" int hcx_pepdb_generate(hcx_pepdb *const *dbs, uint32_t n) {
    if (n == 0) return HCX_SUCCESS;
    if (!dbs) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_pepdb_generate: null");
    hcx_ctx *c = nullptr;
    uint64_t total = 0;
    for (uint32_t j = 0; j < n; j++) {
        if (!dbs[j]) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_pepdb_generate: null database");
        if (c && dbs[j]->ctx != c) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_pepdb_generate: mixed contexts");
        c = dbs[j]->ctx;
        for (uint32_t k = 0; k < j; k++)
            if (dbs[k] == dbs[j]) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_pepdb_generate: a database twice");
        total += (uint64_t)dbs[j]->lclpep + dbs[j]->lclmod;
    }
    if (total > 0xFFFFFFF0ull) return set_error_code(HCX_ERR_INVLD_SIZE, "hcx_pepdb_generate: more than 2^32 forms");
    HCX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    hcx::DeviceBuf tmp, keys, vals, err;
    if (tmp.ensure(sizeof(hcx_pep_entry) * total + 64) || keys.ensure(4 * total + 64) || vals.ensure(4 * total + 64) ||
        err.ensure(64))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_pepdb_generate: device allocation failed");
    HCX_HIP(hipMemsetAsync(err.p, 0, 64, st));
    for (uint32_t j = 0; j < n; j++) {
        hcx_pepdb *db = dbs[j];
        const uint64_t tot = (uint64_t)db->lclpep + db->lclmod;
        if (db->entries.ensure(sizeof(hcx_pep_entry) * tot + 64) || db->mass.ensure(4 * tot + 64))
            return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_pepdb_generate: device allocation failed");
    }
    HCX_HIP(hipEventRecord(c->ev2, st));
    std::vector<uint2> seg;
    uint64_t off = 0;
    for (uint32_t j = 0; j < n; j++) {
        hcx_pepdb *db = dbs[j];
        const uint32_t tot = db->lclpep + db->lclmod;
        hcx_pep_entry *dst = (hcx_pep_entry *)tmp.p + off;
        float *kp = (float *)keys.p + off;
        uint32_t *vp = (uint32_t *)vals.p + off;
        if (db->lclpep)
            hcx::k_lbe_unmod<<<(db->lclpep + 255) / 256, 256, 0, st>>>(
                (const char *)db->seqs.p, db->peplen, db->lclpep, db->cfg.nodes, db->cfg.myid, db->cfg.policy, dst,
                kp, vp, (uint32_t)off);
        if (db->lclmod)
            hcx::k_lbe_mods<<<(db->nseqs + hcx::kModT - 1) / hcx::kModT, hcx::kModT, 0, st>>>(
                (const char *)db->seqs.p, db->nseqs, db->peplen, db->mods, (const uint32_t *)db->var_off.p,
                db->lclpep, db->lclmod, db->cfg.nodes, db->cfg.myid, db->cfg.policy, dst, kp, vp, (uint32_t)off,
                (uint32_t *)err.p);
        HCX_HIP(hipGetLastError());
        if (db->cfg.sort) seg.push_back(make_uint2((uint32_t)off, (uint32_t)(off + tot)));
        off += tot;
    }
    uint32_t he = 0;
    HCX_HIP(hipMemcpyAsync(&he, err.p, 4, hipMemcpyDeviceToHost, st));
    HCX_HIP(hipStreamSynchronize(st));
    if (he) return set_error_code(HCX_ERR_INVLD_MOD, "hcx_pepdb_generate: variant enumeration and count disagree");
    if (!seg.empty())
        if (int rc = hcx::sort_mass_segments(c, (float *)keys.p, (uint32_t *)vals.p, total, seg)) return rc;
    off = 0;
    for (uint32_t j = 0; j < n; j++) {
        hcx_pepdb *db = dbs[j];
        const uint32_t tot = db->lclpep + db->lclmod;
        if (tot) {
            if (db->cfg.sort) {
                hcx::k_lbe_gather<<<(tot + 255) / 256, 256, 0, st>>>(
                    (const hcx_pep_entry *)tmp.p, (const float *)keys.p + off, (const uint32_t *)vals.p + off, tot,
                    (hcx_pep_entry *)db->entries.p, (float *)db->mass.p);
            } else {
                HCX_HIP(hipMemcpyAsync(db->entries.p, (const hcx_pep_entry *)tmp.p + off, sizeof(hcx_pep_entry) * tot,
                                       hipMemcpyDeviceToDevice, st));
                hcx::k_lbe_mass<<<(tot + 255) / 256, 256, 0, st>>>((const hcx_pep_entry *)db->entries.p, tot,
                                                                   (float *)db->mass.p);
            }
            HCX_HIP(hipGetLastError());
        }
        off += tot;
    }
    HCX_HIP(hipEventRecord(c->ev1, st));
    HCX_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev2, c->ev1);
    for (uint32_t j = 0; j < n; j++) {
        dbs[j]->generated = true;
        dbs[j]->generate_ms = ms;
    }
    return HCX_SUCCESS;
}

extern "C" int hcx_pepdb_export(const hcx_pepdb *db, char *seqs, hcx_pep_entry *entries) {
    if (!db) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_pepdb_export: null");
    HCX_HIP(hipSetDevice(db->ctx->device));
    HCX_HIP(hipStreamSynchronize(db->ctx->stream));
    if (seqs && db->nseqs) HCX_HIP(hipMemcpy(seqs, db->seqs.p, (size_t)db->nseqs * db->peplen, hipMemcpyDeviceToHost));
    if (entries) {
        if (!db->generated) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_pepdb_export: not generated");
        const size_t tot = (size_t)db->lclpep + db->lclmod;
        if (tot) HCX_HIP(hipMemcpy(entries, db->entries.p, sizeof(hcx_pep_entry) * tot, hipMemcpyDeviceToHost));
    }
    return HCX_SUCCESS;
}

extern "C" int hcx_pepdb_ms(const hcx_pepdb *db, float *upload_ms, float *count_ms, float *generate_ms) {
    if (!db) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_pepdb_ms: null");
    if (upload_ms) *upload_ms = db->upload_ms;
    if (count_ms) *count_ms = db->count_ms;
    if (generate_ms) *generate_ms = db->generate_ms;
    return HCX_SUCCESS;
}

extern "C" int hcx_pepdb_free(hcx_pepdb *db) {
    if (!db) return HCX_SUCCESS;
    (void)hipSetDevice(db->ctx->device);
    (void)hipStreamSynchronize(db->ctx->stream);
    delete db;
    return HCX_SUCCESS;
}

extern "C" int hcx_index_build_pepdb(hcx_ctx *c, const hcx_pepdb *db, hcx_index **out) {
    if (!c || !db || !out) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_index_build_pepdb: null");
    if (db->ctx != c) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build_pepdb: another context's database");
    if (!db->generated) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_index_build_pepdb: not generated");
    if (!db->cfg.sort)
        return set_error_code(HCX_ERR_INVLD_SORT, "hcx_index_build_pepdb: the query needs entries sorted by mass");
    return hcx::index_build_device(c, db->peplen, (const char *)db->seqs.p, db->nseqs,
                                   (const hcx_pep_entry *)db->entries.p, (const float *)db->mass.p,
                                   db->lclpep + db->lclmod, out);
}

extern "C" int hcx_sort_mass(hcx_ctx *c, float *keys, uint32_t *vals, uint32_t n) {
    if (!c || (n && (!keys || !vals))) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_sort_mass: null");
    for (uint32_t i = 0; i < n; i++)
        if (keys[i] != keys[i]) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_sort_mass: NaN key");
    if (n < 2) return HCX_SUCCESS;
    HCX_HIP(hipSetDevice(c->device));
    hcx::DeviceBuf k, v;
    if (k.ensure(4ull * n + 64) || v.ensure(4ull * n + 64))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_sort_mass: device allocation failed");
    HCX_HIP(hipMemcpyAsync(k.p, keys, 4ull * n, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemcpyAsync(v.p, vals, 4ull * n, hipMemcpyHostToDevice, c->stream));
    std::vector<uint2> seg{make_uint2(0u, n)};
    if (int rc = hcx::sort_mass_segments(c, (float *)k.p, (uint32_t *)v.p, n, seg)) return rc;
    HCX_HIP(hipMemcpy(keys, k.p, 4ull * n, hipMemcpyDeviceToHost));
    HCX_HIP(hipMemcpy(vals, v.p, 4ull * n, hipMemcpyDeviceToHost));
    return HCX_SUCCESS;
}
