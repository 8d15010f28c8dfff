// k_pick.hip -- the Queries producer: MSQuery::pickpeaks (msquery.cpp:509-569)
// over a batch of raw spectra, as MSQuery::extractbatch (msquery.cpp:238-298)
// calls it once per spectrum, and the C ABI around it (include/hcx.h).
//
// One wavefront per spectrum, one launch per size class of spectra (LDS sized
// for the class, so small spectra fit more workgroups per CU).  The
// spectrum's raw (intensity, m/z) pairs are staged in LDS (up to kPickLdsMax
// peaks; larger spectra in a device scratch copy) and sorted with the
// reference's own Hoare quicksort (keyval.h:48-132): the sort is unstable,
// and which of several equal intensities survive the QALEN cut, and in what
// order they are stored, is the order that partition scheme leaves.  Its
// partitions are evaluated in closed form, by the whole wavefront for large
// ranges and for many small ranges at once, one key per lane.  Sub-ranges
// lying wholly below the kept top-QALEN suffix are not sorted further:
// partitions only move elements inside their own range, so the suffix ends
// exactly as in a full sort.  The whole wavefront then scales, counts and
// writes the kept peaks, and computes the batch facts the search needs
// (largest kept count, largest intensity, most peaks of one spectrum within
// 2*dF+1 bins).
//
// Device data per pass: staging rows at ub[s] = sum over t < s of
// min(S_t, QALEN); a device scan of the kept counts gives Queries.idx; a copy
// kernel packs the rows into the dense CSR arrays of the batch, and a small
// reduction the batch facts.
#include <algorithm>
#include <new>
#include <vector>

#include "hcx_internal.h"

struct hcx_rawbatch {
    hcx_ctx *ctx = nullptr;
    int32_t nspec = 0, file_num = 0;
    uint64_t npeaks = 0;
    uint32_t max_len = 0;            // most raw peaks of one spectrum
    bool has_rt = false;
    std::vector<uint32_t> h_idx;     // host copy of the row offsets
    // spectra grouped by size: bin k (k < 3) holds those of at most 256 << k
    // peaks, bin 3 the larger ones; bin_off[k] .. bin_off[k+1] in `order`
    uint32_t bin_off[5] = {};
    hcx::DeviceBuf mz, in, idx, prec, z, rt, order;
    hcx::DeviceBuf big_off, scr;     // scratch rows of spectra above lds_cap
    hcx::DeviceBuf ub, cnt, st, summary, aux, scan;   // per-pass staging
};

namespace hcx {
namespace {

constexpr uint32_t kPickLdsMax = 1024;   // raw peaks of one spectrum staged in LDS
constexpr int kPickBins = 3;             // LDS size classes of 256, 512 and 1024 peaks
static_assert(sizeof(hcx_rawbatch::bin_off) / sizeof(uint32_t) == kPickBins + 2, "one offset per class and the end");
constexpr uint32_t kQalenMax = 4096;
constexpr int kStackMax = 64;            // > log2(2^32): ranges pushed by the one-lane sort

struct PickArgs {
    const uint32_t *order;       // the spectra of this launch, one per workgroup
    const uint32_t *raw_mz, *raw_in, *raw_idx;
    const double *prec_mz;
    const int32_t *z;
    const double *rtime;         // may be null: 0
    const uint32_t *big_off;     // scratch row of a spectrum above lds_cap
    uint32_t *scr_key, *scr_val;
    const uint32_t *ub;          // staging row of each spectrum
    uint32_t *st_moz, *st_in, *cnt;
    float *prec;
    int32_t *chg;
    float *rt;
    int32_t n;
    uint32_t qalen, base_int, min_int;
    uint32_t lds_cap;            // spectra of this launch have at most lds_cap peaks (0: scratch rows)
    uint32_t tmp_words;          // the sort's range list (lds_cap/2 pairs)
};

// keyval.h:50-80: one partition of [l, r] around key[(l+r)/2].  The scans
// stop at the pivot or at a swapped element, never outside [l, r] (the bounds
// in the loop conditions only make that explicit).
__device__ __forceinline__ void partition(uint32_t *key, uint32_t *val, int l, int r, int &io, int &jo) {
    const uint32_t p = key[(l + r) >> 1];
    int i = l, j = r;
    while (i <= j) {
        while (i < r && key[i] < p) i++;
        while (j > l && key[j] > p) j--;
        if (i <= j) {
            const uint32_t tk = key[i], tv = val[i];
            key[i] = key[j];
            val[i] = val[j];
            key[j] = tk;
            val[j] = tv;
            i++;
            j--;
        }
    }
    io = i;
    jo = j;
}

// keyval.h:48-105 on key[0, n), one lane.  Both halves of every partition are
// sorted (keyval.h:82-105), the larger pushed, so the stack stays below
// log2(n) + 1 ranges; ranges ending below `cut` are dropped.
__device__ void keyval_sort_lane(uint32_t *key, uint32_t *val, int n, int cut, int *stack) {
    if (n < 2) return;
    int sp = 0, l = 0, r = n - 1;
    for (;;) {
        int i, j;
        partition(key, val, l, r, i, j);
        const bool go_l = l < j && j >= cut;
        const bool go_r = i < r;
        if (go_l && go_r) {
            if (j - l > r - i) {
                stack[2 * sp] = l;
                stack[2 * sp + 1] = j;
                l = i;
            } else {
                stack[2 * sp] = i;
                stack[2 * sp + 1] = r;
                r = j;
            }
            sp++;
        } else if (go_l) {
            r = j;
        } else if (go_r) {
            l = i;
        } else {
            if (sp == 0) break;
            sp--;
            l = stack[2 * sp];
            r = stack[2 * sp + 1];
        }
    }
}

// Index of the (n+1)-th set bit of m (m has more than n set bits).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int n) {
    int pos = 0;
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (n >= c) {
            n -= c;
            m >>= w;
            pos += w;
        } else {
            m = low;
        }
    }
    return pos;
}

// Per-chunk masks of a cooperative partition (ranges of at most kPickLdsMax).
struct CoopChunks {
    uint64_t ge[kPickLdsMax / 64];
    uint64_t le[kPickLdsMax / 64];
    uint32_t pge[kPickLdsMax / 64];   // ge elements before the chunk
    uint32_t ple[kPickLdsMax / 64];   // le elements before the chunk
};

static_assert(sizeof(CoopChunks) <= kStackMax * 8 && 128 * sizeof(int) <= kStackMax * 8,
              "chunk masks and rank tables share the one-lane sort's stack space");

// Position of the (k+1)-th set element, chunks scanned from the left.
__device__ __forceinline__ int coop_select(const uint64_t *mask, const uint32_t *pre, int nch, int l, int k) {
    int c = 0;
    while (c + 1 < nch && (int)pre[c + 1] <= k) c++;
    return l + 64 * c + nth_set_bit(mask[c], k - (int)pre[c]);
}

// keyval.h:50-80 by the whole wavefront.  With L the ascending positions of
// keys >= pivot and R the descending positions of keys <= pivot, the scans of
// the reference swap L[k] with R[k] exactly while L[k] <= R[k]; that count is
// K = max over x of min(#L in [l, x], #R in [x, r]), and the final scan
// positions follow from L[K-1], R[K-1], L[K], R[K].  Every swapped pair is
// handled by one lane, and pairs never share a position (L[k] == R[k] swaps
// an element with itself).
__device__ void coop_partition(uint32_t *key, uint32_t *val, int l, int r, CoopChunks *cc, int &io, int &jo) {
    const int lane = threadIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t p = key[(l + r) >> 1];
    const int n = r - l + 1;
    const int nch = (n + 63) >> 6;
    int tge = 0, tle = 0;
    for (int c = 0; c < nch; c++) {
        const int x = l + 64 * c + lane;
        const bool in = x <= r;
        const uint32_t k = in ? key[x] : 0u;
        const uint64_t bg = __ballot(in && k >= p), bl = __ballot(in && k <= p);
        if (lane == 0) {
            cc->ge[c] = bg;
            cc->le[c] = bl;
            cc->pge[c] = (uint32_t)tge;
            cc->ple[c] = (uint32_t)tle;
        }
        tge += __popcll(bg);
        tle += __popcll(bl);
    }
    __syncthreads();
    int best = 0;
    for (int c = 0; c < nch; c++) {
        if (l + 64 * c + lane <= r) {
            const int a = (int)cc->pge[c] + __popcll(cc->ge[c] & (lt | (1ull << lane)));
            const int b = tle - (int)cc->ple[c] - __popcll(cc->le[c] & lt);
            best = max(best, min(a, b));
        }
    }
    for (int d = 32; d > 0; d >>= 1) best = max(best, __shfl_xor(best, d, 64));
    const int K = best;
    for (int k = lane; k < K; k += 64) {
        const int xl = coop_select(cc->ge, cc->pge, nch, l, k);
        const int xr = coop_select(cc->le, cc->ple, nch, l, tle - 1 - k);
        if (xl != xr) {
            const uint32_t kl = key[xl], vl = val[xl];
            key[xl] = key[xr];
            val[xl] = val[xr];
            key[xr] = kl;
            val[xr] = vl;
        }
    }
    const int lk1 = coop_select(cc->ge, cc->pge, nch, l, K - 1);
    const int rk1 = coop_select(cc->le, cc->ple, nch, l, tle - K);
    const int i0 = lk1 + 1, j0 = rk1 - 1;
    if (i0 > j0) {
        io = i0;
        jo = j0;
    } else {
        const int lk = K < tge ? coop_select(cc->ge, cc->pge, nch, l, K) : r + 1;
        const int rk = K < tle ? coop_select(cc->le, cc->ple, nch, l, tle - 1 - K) : l - 1;
        io = lk <= j0 ? lk : rk1;
        jo = rk >= i0 ? rk : lk1;
    }
    __syncthreads();
}

// Index of the (n+1)-th highest set bit of m.
__device__ __forceinline__ int nth_set_bit_top(uint64_t m, int n) { return nth_set_bit(m, __popcll(m) - 1 - n); }

// keyval.h:50-105 on ranges of fewer than 64 keys packed into one 64-lane
// window, one key per lane: every range's partitions run at once, level by
// level, until each range is sorted or has dropped below `cut`.  Within a
// range (lanes [sl, sr]), a key >= pivot whose count of keys >= pivot up to
// it is a and whose count of keys <= pivot from it on is b is swapped exactly
// when b >= a, with the a-th key <= pivot from the right; a key <= pivot
// symmetrically (a >= b, the b-th key >= pivot from the left).  The final scan
// positions follow as in coop_partition.  Rank tables in LDS (posL: lane of
// the k-th key >= pivot, posR: lane of the k-th key <= pivot from the right,
// both indexed from the range's first lane) give the partners.  Ranges:
// list[first, first+count).
__device__ void window_sort(uint32_t *key, uint32_t *val, const int *list, int first, int count, int cut,
                            int *posL, int *posR) {
    const int lane = threadIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull;
    int start = 0, sl = 0, sr = -1, base = 0;
    bool in = false, act = false;
    for (int g = 0; g < count; g++) {
        const int l = list[2 * (first + g)], r = list[2 * (first + g) + 1];
        const int n = r - l + 1;
        if (lane >= start && lane < start + n) {
            sl = start;
            sr = start + n - 1;
            base = l - start;
            in = true;
            act = n >= 2;
        }
        start += n;
    }
    uint32_t k = 0, v = 0;
    if (in) {
        k = key[lane + base];
        v = val[lane + base];
    }
    while (__ballot(act)) {
        const uint32_t p = (uint32_t)__shfl((int)k, (sl + sr) >> 1, 64);
        const uint64_t GE = __ballot(act && k >= p), LE = __ballot(act && k <= p);
        const uint64_t seg = act ? ((sr == 63 ? ~0ull : (1ull << (sr + 1)) - 1ull) & ~((1ull << sl) - 1ull)) : 0ull;
        const uint64_t ge = GE & seg, le = LE & seg;
        const int a = __popcll(ge & (lt | (1ull << lane)));
        const int b = __popcll(le & ~lt);
        const bool isge = (GE >> lane) & 1ull, isle = (LE >> lane) & 1ull;
        if (isge) posL[sl + a - 1] = lane;
        if (isle) posR[sl + b - 1] = lane;
        const bool swg = act && isge && b >= a;
        const uint64_t SW = __ballot(swg);
        __syncthreads();
        int src = lane, I = 0, J = 0;
        if (act) {
            if (swg) src = posR[sl + a - 1];
            else if (isle && a >= b) src = posL[sl + b - 1];
            const int K = __popcll(SW & seg);
            const int lk1 = posL[sl + K - 1], rk1 = posR[sl + K - 1];
            const int i0 = lk1 + 1, j0 = rk1 - 1;
            if (i0 > j0) {
                I = i0;
                J = j0;
            } else {
                const int lk = K < __popcll(ge) ? posL[sl + K] : sr + 1;
                const int rk = K < __popcll(le) ? posR[sl + K] : sl - 1;
                I = lk <= j0 ? lk : rk1;
                J = rk >= i0 ? rk : lk1;
            }
        }
        __syncthreads();
        k = (uint32_t)__shfl((int)k, src, 64);
        v = (uint32_t)__shfl((int)v, src, 64);
        if (act) {
            // keyval.h:82-105: [l, j] if l < j (and not below the cut), [i, r] if i < r
            if (lane <= J) {
                if (sl < J && J + base >= cut) sr = J;
                else act = false;
            } else if (lane >= I) {
                if (I < sr) sl = I;
                else act = false;
            } else {
                act = false;
            }
        }
    }
    if (in) {
        key[lane + base] = k;
        val[lane + base] = v;
    }
}

// The same sort with the whole wavefront: pending ranges sit in an LDS list.
// A range of at least 64 keys among the top 64 entries is partitioned by all
// lanes together (its halves go back on the list); otherwise ranges from the
// top of the list are packed into a 64-lane window and sorted there
// completely.  Ranges never overlap, so the order in which they are processed
// does not change the result.  Pending ranges are disjoint and at least two
// long, so the list never holds more than n/2.
__device__ void keyval_sort_wave(uint32_t *key, uint32_t *val, int n, int cut, int *list, CoopChunks *cc) {
    const int lane = threadIdx.x;
    int cnt = 0;
    if (n >= 2) {
        if (lane == 0) {
            list[0] = 0;
            list[1] = n - 1;
        }
        cnt = 1;
    }
    __syncthreads();
    while (cnt > 0) {
        const int take = min(cnt, 64);
        const int base = cnt - take;
        int l = 0, r = -1;
        if (lane < take) {
            l = list[2 * (base + lane)];
            r = list[2 * (base + lane) + 1];
        }
        const uint64_t big = __ballot(lane < take && r - l + 1 >= 64);
        __syncthreads();
        if (big) {
            const int src = __ffsll((unsigned long long)big) - 1;
            const int cl = __shfl(l, src, 64), cr = __shfl(r, src, 64);
            if (lane == 0) {            // the top entry takes the chosen one's slot
                list[2 * (base + src)] = list[2 * (cnt - 1)];
                list[2 * (base + src) + 1] = list[2 * (cnt - 1) + 1];
            }
            cnt--;
            int i, j;
            coop_partition(key, val, cl, cr, cc, i, j);
            if (lane == 0) {
                int w = cnt;
                if (cl < j && j >= cut) {
                    list[2 * w] = cl;
                    list[2 * w + 1] = j;
                    w++;
                }
                if (i < cr) {
                    list[2 * w] = i;
                    list[2 * w + 1] = cr;
                }
            }
            cnt += (cl < j && j >= cut ? 1 : 0) + (i < cr ? 1 : 0);
            __syncthreads();
            continue;
        }
        // every top range is below 64 keys: pack them from the top (at most
        // 32 of them, all within the 64 just checked)
        int m = 0, tot = 0;
        while (m < cnt) {
            const int e = cnt - 1 - m;
            const int sz = list[2 * e + 1] - list[2 * e] + 1;
            if (tot + sz > 64) break;
            tot += sz;
            m++;
        }
        window_sort(key, val, list, cnt - m, m, cut, reinterpret_cast<int *>(cc), reinterpret_cast<int *>(cc) + 64);
        cnt -= m;
        __syncthreads();
    }
}

// msquery.cpp:537: dIntArr[j] *= factor, a uint times a double truncated back
// to uint.  0 * inf (a spectrum whose top intensity is 0) is undefined in C;
// x86-64 converts it to 0, and so does this.
__device__ __forceinline__ uint32_t scale_int(uint32_t k, double factor) {
    const double v = (double)k * factor;
    return v == v ? (uint32_t)v : 0u;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int d = 32; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
    return v;
}

// One spectrum: key/val are its raw (intensity, m/z) arrays, in LDS or in the
// scratch rows (the two call sites let the compiler use LDS instructions for
// the first); tmp (the sort's range list, then the kept bins) and stack are
// LDS.  Spectra in LDS are sorted by the whole wavefront, larger ones by one
// lane (a range list for them would not fit).
template <bool IN_LDS>
__device__ __forceinline__ void pick_one(const PickArgs &a, int s, int S, uint32_t b, uint32_t *key, uint32_t *val,
                                         uint32_t *tmp, int *stack) {
    const int lane = threadIdx.x;
    for (int t = lane; t < S; t += 64) {
        key[t] = a.raw_in[b + t];
        val[t] = a.raw_mz[b + t];
    }
    __syncthreads();
    // msquery.cpp:519, only the part the kept suffix depends on
    const int cut = S > (int)a.qalen ? S - (int)a.qalen : 0;
    if (IN_LDS) {
        keyval_sort_wave(key, val, S, cut, reinterpret_cast<int *>(tmp), reinterpret_cast<CoopChunks *>(stack));
    } else {
        if (lane == 0) keyval_sort_lane(key, val, S, cut, stack);
        __syncthreads();
    }

    // msquery.cpp:524-545: factor, top -> base_int, count of scaled >= min_int
    uint32_t speclen = 0;
    double factor = 0.0;
    if (S > 0) {
        factor = (double)a.base_int / (double)key[S - 1];
        uint32_t c = 0;
        for (int j0 = cut; j0 <= S - 2; j0 += 64) {
            const int j = j0 + lane;
            const bool f = j <= S - 2 && scale_int(key[j], factor) >= a.min_int;
            c += (uint32_t)__popcll(__ballot(f));
        }
        speclen = 1 + c;
    }
    // msquery.cpp:548-562: the last min(speclen, QALEN) peaks, ascending intensity
    const int keep = (int)min(speclen, a.qalen);
    const uint32_t o = a.ub[s];
    for (int t = lane; t < keep; t += 64) {
        const int pos = S - keep + t;
        a.st_in[o + t] = pos == S - 1 ? a.base_int : scale_int(key[pos], factor);
        a.st_moz[o + t] = val[pos];
    }
    if (lane == 0) {
        a.cnt[s] = (uint32_t)keep;
        // msquery.cpp:515-517
        a.prec[s] = (float)a.prec_mz[s];
        const int32_t zz = a.z[s];
        a.chg[s] = 1 > zz ? 1 : zz;
        const double rt = a.rtime ? a.rtime[s] : 0.0;
        a.rt[s] = (float)(0.0 > rt ? 0.0 : rt);
        if (s == 0) a.cnt[a.n] = 0;   // the scan's last term: idx[n] = total
    }
}

__global__ __launch_bounds__(64) void k_pickpeaks(PickArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint32_t *lds = reinterpret_cast<uint32_t *>(smem_raw);
    const int s = (int)a.order[blockIdx.x];
    const uint32_t b = a.raw_idx[s];
    const int S = (int)(a.raw_idx[s + 1] - b);
    uint32_t *tmp = lds + 2 * a.lds_cap;
    int *stack = reinterpret_cast<int *>(tmp + a.tmp_words);
    if ((uint32_t)S <= a.lds_cap)
        pick_one<true>(a, s, S, b, lds, lds + a.lds_cap, tmp, stack);
    else
        pick_one<false>(a, s, S, b, a.scr_key + a.big_off[s], a.scr_val + a.big_off[s], tmp, stack);
}

// Batch facts: the largest kept count, intensity and multiplicity.
__global__ __launch_bounds__(256) void k_pick_max(const uint32_t *cnt, const uint32_t *imax_s, const uint32_t *mult_s,
                                                  int n, uint32_t *summary) {
    uint32_t c = 0, m = 0, u = 0;
    for (int s = blockIdx.x * 256 + threadIdx.x; s < n; s += gridDim.x * 256) {
        c = max(c, cnt[s]);
        m = max(m, imax_s[s]);
        u = max(u, mult_s[s]);
    }
    c = wave_max_u32(c);
    m = wave_max_u32(m);
    u = wave_max_u32(u);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&summary[0], c);
        atomicMax(&summary[1], m);
        atomicMax(&summary[2], u);
    }
}

// staging rows -> dense CSR, one wavefront per spectrum
__global__ __launch_bounds__(256) void k_pick_compact(const uint32_t *ub, const uint32_t *cnt, const uint32_t *idx,
                                                      const uint32_t *st_moz, const uint32_t *st_in, uint32_t *moz,
                                                      uint32_t *in, int n) {
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= n) return;
    const uint32_t o = ub[s], d = idx[s], c = cnt[s];
    for (uint32_t t = lane; t < c; t += 64) {
        moz[d + t] = st_moz[o + t];
        in[d + t] = st_in[o + t];
    }
}

// Batch facts of a Queries batch in HBM (uploaded as it is, or made by
// pickpeaks): per spectrum its peak count, its largest intensity, the most
// peaks within span + 1 consecutive bins and the largest intensity sum within
// span + 1 bins (saturating at 2^32 - 1), by a broadcast count over its peaks
// (for each peak, the peaks in [bin, bin + span]).  One wavefront per
// spectrum; bins and intensities are staged in LDS when the spectrum has at
// most `cap` peaks (a multiple of 4), else read from HBM.
// STAGED: bins and ints are 16-byte aligned LDS arrays, read four at a time
template <bool STAGED>
__device__ __forceinline__ void window_facts(const uint32_t *bins, const uint32_t *ints, uint32_t P, uint32_t span,
                                             uint32_t lane, uint32_t *mult, uint32_t *wsum) {
    uint32_t m = 0;
    unsigned long long wm = 0;
    const uint32_t P4 = STAGED ? P & ~3u : 0u;
    for (uint32_t t0 = 0; t0 < P; t0 += 64) {
        const uint32_t t = t0 + lane;
        const uint32_t mx = t < P ? bins[t] : 0u;
        uint32_t c = 0;
        unsigned long long w = 0;
        auto add = [&](uint32_t v, uint32_t i) {
            const bool in = v >= mx && v - mx <= span;
            c += in ? 1u : 0u;
            w += in ? i : 0u;
        };
        for (uint32_t y = 0; y < P4; y += 4) {
            const uint4 v = *reinterpret_cast<const uint4 *>(bins + y);
            const uint4 i = *reinterpret_cast<const uint4 *>(ints + y);
            add(v.x, i.x);
            add(v.y, i.y);
            add(v.z, i.z);
            add(v.w, i.w);
        }
        for (uint32_t y = P4; y < P; y++) add(bins[y], ints[y]);
        if (t < P) {
            m = max(m, c);
            wm = max(wm, w);
        }
    }
    *mult = m;
    *wsum = (uint32_t)min(wm, 0xFFFFFFFFull);
}

__global__ __launch_bounds__(64) void k_batch_facts(const uint32_t *moz, const uint32_t *inten, const uint32_t *idx,
                                                    uint32_t span, uint32_t cap, uint32_t *cnt_s, uint32_t *imax_s,
                                                    uint32_t *mult_s, uint32_t *wsum_s) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint32_t *tb = reinterpret_cast<uint32_t *>(smem_raw);
    uint32_t *ti = tb + cap;
    const int s = (int)blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t b = idx[s], P = idx[s + 1] - b;
    const bool staged = P <= cap;
    uint32_t imax = 0;
    for (uint32_t t = lane; t < P; t += 64) {
        const uint32_t in = inten[b + t];
        imax = max(imax, in);
        if (staged) {
            tb[t] = moz[b + t];
            ti[t] = in;
        }
    }
    __syncthreads();
    uint32_t mult, ws;
    if (staged) window_facts<true>(tb, ti, P, span, lane, &mult, &ws);
    else window_facts<false>(moz + b, inten + b, P, span, lane, &mult, &ws);
    imax = wave_max_u32(imax);
    mult = wave_max_u32(mult);
    ws = wave_max_u32(ws);
    if (lane == 0) {
        cnt_s[s] = P;
        imax_s[s] = imax;
        mult_s[s] = mult;
        wsum_s[s] = ws;
    }
}

}  // namespace

// The batch facts of b (moz, inten, idx resident): per spectrum into b->mult
// and b->wsum (allocated by the caller), and the maxima of the peak counts,
// intensities and multiplicities into summary[0..2] (16 bytes, zeroed here).
// tmp: 2 * nspec words.  maxp bounds the largest spectrum.  Enqueues only.
int batch_facts_launch(hcx_ctx *c, hcx_qbatch *b, uint32_t maxp, uint32_t *tmp, uint32_t *summary) {
    const int n = b->nspec;
    if (n <= 0) return HCX_SUCCESS;
    hipStream_t st = c->stream;
    // a multiple of 4, so the intensities after the bins stay 16-byte aligned
    const uint32_t cap = (std::min<uint32_t>(maxp, 4096) + 3u) & ~3u;
    HCX_HIP(hipMemsetAsync(summary, 0, 16, st));
    k_batch_facts<<<(unsigned)n, 64, 8u * cap + 16, st>>>(
        (const uint32_t *)b->moz.p, (const uint32_t *)b->inten.p, (const uint32_t *)b->idx.p, 2 * c->params.dF, cap,
        tmp, tmp + n, (uint32_t *)b->mult.p, (uint32_t *)b->wsum.p);
    HCX_HIP(hipGetLastError());
    k_pick_max<<<(unsigned)std::min<int32_t>((n + 255) / 256, 1024), 256, 0, st>>>(tmp, tmp + n, (const uint32_t *)b->mult.p,
                                                                                  n, summary);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int batch_facts(hcx_ctx *c, hcx_qbatch *b, uint32_t maxp) {
    const int n = b->nspec;
    b->max_peaks = b->max_int = b->max_mult = 0;
    if (b->mult.ensure(4ull * std::max(n, 0) + 16) || b->wsum.ensure(4ull * std::max(n, 0) + 16))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "batch facts: device allocation failed");
    if (n <= 0) return HCX_SUCCESS;
    DeviceBuf tmp, summary;
    if (tmp.ensure(8ull * n + 16) || summary.ensure(64))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "batch facts: device allocation failed");
    if (int rc = batch_facts_launch(c, b, maxp, (uint32_t *)tmp.p, (uint32_t *)summary.p)) return rc;
    uint32_t h[4] = {0, 0, 0, 0};
    HCX_HIP(hipMemcpyAsync(h, summary.p, 12, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    b->max_peaks = h[0];
    b->max_int = h[1];
    b->max_mult = h[2];
    return HCX_SUCCESS;
}

}  // namespace hcx

using hcx::set_error;
using hcx::set_error_code;

extern "C" int hcx_raw_upload(hcx_ctx *c, const hcx_raw_spectra_view *raw, hcx_rawbatch **out) {
    if (!c || !raw || !out) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_raw_upload: null argument");
    if (raw->num_specs < 0) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_raw_upload: negative count");
    const int32_t n = raw->num_specs;
    if (n > 0 && (!raw->idx || !raw->prec_mz || !raw->z))
        return set_error_code(HCX_ERR_INVLD_PTR, "hcx_raw_upload: missing arrays");
    if (n > 0 && raw->idx[0] != 0) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_raw_upload: idx[0] != 0");
    uint32_t maxS = 0;
    uint64_t nbig = 0;
    std::vector<uint32_t> big(n > 0 ? n : 1, 0), order(n > 0 ? n : 1, 0), binof(n > 0 ? n : 1, 0);
    uint32_t bin_off[hcx::kPickBins + 2] = {};
    for (int32_t s = 0; s < n; s++) {
        if (raw->idx[s + 1] < raw->idx[s]) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_raw_upload: idx not monotone");
        const uint32_t S = raw->idx[s + 1] - raw->idx[s];
        maxS = std::max(maxS, S);
        uint32_t k = 0;
        while (k < (uint32_t)hcx::kPickBins && S > (256u << k)) k++;
        binof[s] = k;
        bin_off[k + 1]++;
        if (S > hcx::kPickLdsMax) {
            big[s] = (uint32_t)nbig;
            nbig += S;
        }
    }
    for (int k = 0; k <= hcx::kPickBins; k++) bin_off[k + 1] += bin_off[k];
    {
        uint32_t fill[hcx::kPickBins + 2];
        std::copy(bin_off, bin_off + hcx::kPickBins + 2, fill);
        for (int32_t s = 0; s < n; s++) order[fill[binof[s]]++] = (uint32_t)s;
    }
    const uint64_t npk = n > 0 ? raw->idx[n] : 0;
    if (npk && (!raw->mz || !raw->intensity)) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_raw_upload: missing peaks");
    HCX_HIP(hipSetDevice(c->device));
    hcx_rawbatch *b = new (std::nothrow) hcx_rawbatch;
    if (!b) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_raw_upload: out of memory");
    b->ctx = c;
    b->nspec = n;
    b->file_num = raw->file_num;
    b->npeaks = npk;
    b->max_len = maxS;
    std::copy(bin_off, bin_off + hcx::kPickBins + 2, b->bin_off);
    b->has_rt = raw->rtime != nullptr;
    b->h_idx.assign(raw->idx, raw->idx + (n > 0 ? n + 1 : 0));
    if (b->mz.ensure(4 * npk + 16) || b->in.ensure(4 * npk + 16) || b->idx.ensure(4ull * (n + 1)) ||
        b->prec.ensure(8ull * n + 16) || b->z.ensure(4ull * n + 16) || b->rt.ensure(8ull * n + 16) ||
        b->big_off.ensure(4ull * n + 16) || b->scr.ensure(8 * nbig + 16) || b->order.ensure(4ull * n + 16)) {
        delete b;
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_raw_upload: device allocation failed");
    }
    auto up = [&](hcx::DeviceBuf &d, const void *h, size_t bytes) {
        if (!bytes || !h) return hipSuccess;
        return hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c->stream);
    };
    if (up(b->mz, raw->mz, 4 * npk) || up(b->in, raw->intensity, 4 * npk) ||
        up(b->idx, raw->idx, 4ull * (n > 0 ? n + 1 : 0)) || up(b->prec, raw->prec_mz, 8ull * n) ||
        up(b->z, raw->z, 4ull * n) || up(b->rt, raw->rtime, 8ull * n) || up(b->big_off, big.data(), 4ull * n) ||
        up(b->order, order.data(), 4ull * n) ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        delete b;
        return set_error("hcx_raw_upload: copy failed");
    }
    *out = b;
    return HCX_SUCCESS;
}

extern "C" int hcx_raw_free(hcx_rawbatch *rb) {
    if (!rb) return HCX_SUCCESS;
    (void)hipSetDevice(rb->ctx->device);
    (void)hipStreamSynchronize(rb->ctx->stream);
    delete rb;
    return HCX_SUCCESS;
}

extern "C" int hcx_queries_pick(hcx_ctx *c, hcx_rawbatch *rb, const hcx_pick_params *pp, hcx_qbatch **out,
                                float *kernel_ms, float *total_ms) {
    if (!c || !rb || !pp || !out) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_queries_pick: null argument");
    if (rb->ctx != c) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_queries_pick: raw batch of another context");
    if (pp->qalen == 0 || pp->qalen > hcx::kQalenMax)
        return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_queries_pick: qalen must be 1..4096");
    const int32_t n = rb->nspec;
    // staging rows: min(S, qalen) per spectrum
    std::vector<uint32_t> ub(n > 0 ? n : 1, 0);
    uint64_t tot = 0;
    for (int32_t s = 0; s < n; s++) {
        ub[s] = (uint32_t)tot;
        tot += std::min<uint64_t>(rb->h_idx[s + 1] - rb->h_idx[s], pp->qalen);
    }
    if (tot > 0xFFFFFFFFull) return set_error_code(HCX_ERR_INVLD_PARAM, "hcx_queries_pick: more than 2^32 kept peaks");
    HCX_HIP(hipSetDevice(c->device));
    hcx_qbatch *b = new (std::nothrow) hcx_qbatch;
    if (!b) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_queries_pick: out of memory");
    b->ctx = c;
    b->nspec = n;
    b->file_num = rb->file_num;
    if (b->moz.ensure(4 * tot + 16) || b->inten.ensure(4 * tot + 16) || b->idx.ensure(4ull * (n + 1)) ||
        b->prec.ensure(4ull * n + 16) || b->chg.ensure(4ull * n + 16) || b->rt.ensure(4ull * n + 16) ||
        b->mult.ensure(4ull * n + 16) || b->wsum.ensure(4ull * n + 16) || rb->ub.ensure(4ull * n + 16) || rb->cnt.ensure(4ull * n + 16) || rb->st.ensure(8 * tot + 16) ||
        rb->summary.ensure(16) || rb->aux.ensure(8ull * n + 16) ||
        rb->scan.ensure(4 * hcx::scan_tmp_words((uint64_t)n + 1) + 64)) {
        delete b;
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "hcx_queries_pick: device allocation failed");
    }
    hipStream_t st = c->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    hipEvent_t ef = nullptr, ej0 = nullptr, ej1 = nullptr;   // fork / joins of the class streams
    auto drop = [&]() {
        for (hipEvent_t e : {e0, e1, e2, e3, ef, ej0, ej1})
            if (e) (void)hipEventDestroy(e);
    };
    auto fail = [&](const char *msg) {
        // nothing launched may still use the batch's buffers
        (void)hipStreamSynchronize(st);
        for (hipStream_t x : c->aux)
            if (x) (void)hipStreamSynchronize(x);
        drop();
        delete b;
        return set_error(msg);
    };
    if (hipEventCreate(&e0) || hipEventCreate(&e1) || hipEventCreate(&e2) || hipEventCreate(&e3))
        return fail("hcx_queries_pick: event creation failed");
    uint32_t summary[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(rb->ub.p, ub.data(), 4ull * n, hipMemcpyHostToDevice, st) ||
        hipMemsetAsync(rb->summary.p, 0, 16, st) || hipEventRecord(e2, st))
        return fail("hcx_queries_pick: set-up failed");
    hcx::PickArgs a{};
    a.raw_mz = (const uint32_t *)rb->mz.p;
    a.raw_in = (const uint32_t *)rb->in.p;
    a.raw_idx = (const uint32_t *)rb->idx.p;
    a.prec_mz = (const double *)rb->prec.p;
    a.z = (const int32_t *)rb->z.p;
    a.rtime = rb->has_rt ? (const double *)rb->rt.p : nullptr;
    a.big_off = (const uint32_t *)rb->big_off.p;
    a.scr_key = (uint32_t *)rb->scr.p;
    a.scr_val = (uint32_t *)rb->scr.p + rb->scr.bytes / 8;
    a.ub = (const uint32_t *)rb->ub.p;
    a.st_moz = (uint32_t *)rb->st.p;
    a.st_in = (uint32_t *)rb->st.p + tot;
    a.cnt = (uint32_t *)rb->cnt.p;
    a.prec = (float *)b->prec.p;
    a.chg = (int32_t *)b->chg.p;
    a.rt = (float *)b->rt.p;
    a.n = n;
    a.qalen = pp->qalen;
    a.base_int = pp->base_int;
    a.min_int = pp->min_int;
    if (hipEventRecord(e0, st)) return fail("hcx_queries_pick: event record failed");
    // one launch per size class, each with the LDS its largest spectrum needs
    // (small spectra then fit more workgroups per CU).  The classes are
    // independent: the second and third run on two auxiliary streams, so one
    // class's tail overlaps the others' work.
    for (int i = 0; i < 2; i++)
        if (!c->aux[i] && hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking) != hipSuccess) {
            c->aux[i] = nullptr;
            return fail("hcx_queries_pick: stream creation failed");
        }
    if (hipEventCreateWithFlags(&ef, hipEventDisableTiming) || hipEventCreateWithFlags(&ej0, hipEventDisableTiming) ||
        hipEventCreateWithFlags(&ej1, hipEventDisableTiming) || hipEventRecord(ef, st) ||
        hipStreamWaitEvent(c->aux[0], ef, 0) || hipStreamWaitEvent(c->aux[1], ef, 0))
        return fail("hcx_queries_pick: stream set-up failed");
    for (int k = 0; k <= hcx::kPickBins; k++) {
        const uint32_t cnt = rb->bin_off[k + 1] - rb->bin_off[k];
        if (!cnt) continue;
        a.order = (const uint32_t *)rb->order.p + rb->bin_off[k];
        a.lds_cap = k < hcx::kPickBins ? 256u << k : 0u;
        // the sort's range list (a multiple of 4 words: the chunk masks after
        // it stay 16-byte aligned); the one-lane sort of the scratch rows
        // needs none
        a.tmp_words = k < hcx::kPickBins ? a.lds_cap : 0u;
        const size_t lds = (2ull * a.lds_cap + a.tmp_words) * 4 + hcx::kStackMax * 8;
        hipStream_t sk = (k == 1 || k == 2) ? c->aux[k - 1] : st;
        hcx::k_pickpeaks<<<cnt, 64, lds, sk>>>(a);
    }
    if (hipGetLastError() || hipEventRecord(ej0, c->aux[0]) || hipEventRecord(ej1, c->aux[1]) ||
        hipStreamWaitEvent(st, ej0, 0) || hipStreamWaitEvent(st, ej1, 0))
        return fail("hcx_queries_pick: normalisation kernel launch failed");
    if (hipGetLastError() || hipEventRecord(e1, st)) return fail("hcx_queries_pick: normalisation kernel launch failed");
    // Queries.idx = exclusive scan of the kept counts and a trailing 0
    // (idx[s+1] = idx[s] + speclen, msquery.cpp:548-549)
    if (n == 0) {
        if (hipMemsetAsync(b->idx.p, 0, 4, st)) return fail("hcx_queries_pick: set-up failed");
    } else if (hcx::scan_exclusive_u32(st, (const uint32_t *)rb->cnt.p, (uint32_t *)b->idx.p, (uint64_t)n + 1,
                                       (uint32_t *)rb->scan.p) != HCX_SUCCESS) {
        return fail("hcx_queries_pick: scan failed");
    }
    if (n > 0)
        hcx::k_pick_compact<<<(unsigned)((n + 3) / 4), 256, 0, st>>>(
            (const uint32_t *)rb->ub.p, (const uint32_t *)rb->cnt.p, (const uint32_t *)b->idx.p, a.st_moz, a.st_in,
            (uint32_t *)b->moz.p, (uint32_t *)b->inten.p, n);
    if (hipGetLastError()) return fail("hcx_queries_pick: launch failed");
    // the batch facts the search sizes its exact scan by, from the batch itself
    if (hcx::batch_facts_launch(c, b, std::min<uint32_t>(pp->qalen, rb->max_len), (uint32_t *)rb->aux.p,
                                (uint32_t *)rb->summary.p) != HCX_SUCCESS ||
        hipEventRecord(e3, st))
        return fail("hcx_queries_pick: launch failed");
    uint32_t npk = 0;
    if (hipMemcpyAsync(summary, rb->summary.p, 12, hipMemcpyDeviceToHost, st) ||
        hipMemcpyAsync(&npk, (const uint32_t *)b->idx.p + n, 4, hipMemcpyDeviceToHost, st) ||
        hipStreamSynchronize(st))
        return fail("hcx_queries_pick: device pass failed");
    float km = 0, tm = 0;
    if (hipEventElapsedTime(&km, e0, e1) || hipEventElapsedTime(&tm, e2, e3))
        return fail("hcx_queries_pick: timing failed");
    drop();
    if (kernel_ms) *kernel_ms = km;
    if (total_ms) *total_ms = tm;
    b->npeaks = npk;
    b->max_peaks = summary[0];
    b->max_int = summary[1];
    b->max_mult = summary[2];
    *out = b;
    return HCX_SUCCESS;
}

extern "C" int hcx_queries_pickpeaks(hcx_ctx *c, const hcx_raw_spectra_view *raw, const hcx_pick_params *pp,
                                     hcx_qbatch **out) {
    hcx_rawbatch *rb = nullptr;
    int rc = hcx_raw_upload(c, raw, &rb);
    if (rc != HCX_SUCCESS) return rc;
    rc = hcx_queries_pick(c, rb, pp, out, nullptr, nullptr);
    hcx_raw_free(rb);
    return rc;
}

extern "C" int hcx_queries_info(const hcx_qbatch *qb, int32_t *num_specs, uint64_t *npeaks) {
    if (!qb) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_queries_info: null batch");
    if (num_specs) *num_specs = qb->nspec;
    if (npeaks) *npeaks = qb->npeaks;
    return HCX_SUCCESS;
}

extern "C" int hcx_queries_facts(const hcx_qbatch *qb, uint32_t *max_peaks, uint32_t *max_int, uint32_t *max_mult,
                                 uint32_t *mult, uint32_t *wsum) {
    if (!qb) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_queries_facts: null batch");
    if (max_peaks) *max_peaks = qb->max_peaks;
    if (max_int) *max_int = qb->max_int;
    if (max_mult) *max_mult = qb->max_mult;
    if (mult && qb->nspec > 0) {
        HCX_HIP(hipSetDevice(qb->ctx->device));
        HCX_HIP(hipStreamSynchronize(qb->ctx->stream));
        HCX_HIP(hipMemcpy(mult, qb->mult.p, 4ull * qb->nspec, hipMemcpyDeviceToHost));
    }
    if (wsum && qb->nspec > 0) {
        HCX_HIP(hipSetDevice(qb->ctx->device));
        HCX_HIP(hipStreamSynchronize(qb->ctx->stream));
        HCX_HIP(hipMemcpy(wsum, qb->wsum.p, 4ull * qb->nspec, hipMemcpyDeviceToHost));
    }
    return HCX_SUCCESS;
}

extern "C" int hcx_queries_export(const hcx_qbatch *qb, uint32_t *moz, uint32_t *intensity, uint32_t *idx,
                                  float *precurse, int32_t *charges, float *rtimes) {
    if (!qb) return set_error_code(HCX_ERR_INVLD_PTR, "hcx_queries_export: null batch");
    hcx_ctx *c = qb->ctx;
    HCX_HIP(hipSetDevice(c->device));
    const uint64_t n = (uint64_t)qb->nspec, k = qb->npeaks;
    auto down = [&](void *h, const hcx::DeviceBuf &d, size_t bytes) {
        if (!h || !bytes) return hipSuccess;
        return hipMemcpyAsync(h, d.p, bytes, hipMemcpyDeviceToHost, c->stream);
    };
    if (down(moz, qb->moz, 4 * k) || down(intensity, qb->inten, 4 * k) || down(idx, qb->idx, 4 * (n + 1)) ||
        down(precurse, qb->prec, 4 * n) || down(charges, qb->chg, 4 * n) || down(rtimes, qb->rt, 4 * n) ||
        hipStreamSynchronize(c->stream))
        return set_error("hcx_queries_export: copy failed");
    return HCX_SUCCESS;
}
