// k_query.hip -- shared-peak scan, hyperscore, top hit and tail fit on the GPU.
//
// DSLIM_QuerySpectrum (dslim_query.cpp:515-842): for every spectrum, every
// peptide length and every peak bin q-dF..q+dF, the iA entries of that bin
// whose ion id lies inside the precursor window add one b- or y-count and the
// peak intensity to their peptide's counters; every peptide with
// bc+yc >= min_shp then gets the hyperscore
//     hs = (float)(lgf[bc] + lgf[yc] + log10(1+ibc) + log10(1+iyc) - 6)
// and, if hs > 0, enters the top-K heap and the 0.1-wide score histogram.
//
// Layout of one workgroup's LDS: the window of one length is tiled into T
// peptides; each peptide has two packed u64 counters (b, y):
//     count << 48 | intensity sum            (one ds_add_u64 per iA entry)
// Every (peak, bin) probe keeps a cursor into its iA bin range, so each iA
// entry is read once no matter how many tiles the window needs.
//
// Top hit: the reference keeps a min-heap (minheap.h) whose getMax() breaks
// ties by heap position.  A unique maximum is the heap's maximum whatever the
// insertion order, so the scan only tracks (max, #ties); spectra whose
// maximum is tied are re-run by the replay kernel, which inserts every
// candidate into an exact copy of the heap in the reference order (length,
// then peptide id).
//
// Tail fit (expeRT.cpp:572-718) runs one wavefront per spectrum in FP64; the
// prefix sums are of integer counts (exact in any order), the regression sums
// run on one lane in the reference order.
#include <cfloat>

#include "hcx_internal.h"

namespace hcx {

void make_divider(uint32_t d, uint32_t *m, uint32_t *s) {
    // smallest l with 2^l >= d; m = floor(2^32 * (2^l - d) / d) + 1
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    uint64_t mm = ((((uint64_t)1 << l) - d) << 32) / d + 1;
    *m = (uint32_t)mm;
    *s = l;
}

namespace {

constexpr int kBT = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr unsigned long long kCnt = 1ull << 48;
constexpr unsigned long long kSumMask = kCnt - 1;

__device__ inline uint32_t udiv(uint32_t n, uint32_t m, uint32_t s) {
    if (s == 0) return n;   // d == 1
    uint32_t t = __umulhi(m, n);
    return (t + ((n - t) >> 1)) >> (s - 1);
}

struct Cell {
    float hs;
    int32_t psid;
    uint32_t ixx, bc, yc, speclen;
};

// first index in [lo, hi) with mass >= v (v > 0 handled by the directory)
__device__ inline uint32_t mass_lb(const float *mass, uint32_t lo, uint32_t hi, float v) {
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (mass[m] < v) lo = m + 1;
        else hi = m;
    }
    return lo;
}
__device__ inline uint32_t mass_ub(const float *mass, uint32_t lo, uint32_t hi, float v) {
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (v < mass[m]) hi = m;
        else lo = m + 1;
    }
    return lo;
}

// Directory-narrowed range for a value v: [mdir[d], mdir[d+1]] with d = floor(v).
__device__ inline void dir_range(const uint32_t *mdir, uint32_t D, uint32_t n, float v, uint32_t *lo,
                                 uint32_t *hi) {
    if (!(v > 0.0f)) {
        *lo = 0;
        *hi = (D > 1) ? mdir[1] : n;
        return;
    }
    uint32_t d = (v >= (float)(D - 1)) ? D - 1 : (uint32_t)v;
    *lo = mdir[d];
    *hi = (d + 1 < D) ? mdir[d + 1] : n;
}

// dslim_query.cpp:854-916 as lower/upper bounds on the float masses
__device__ inline void window(const LenDesc &L, float prec, double dM, uint32_t D, int32_t *wlo, int32_t *whi) {
    *wlo = 1;
    *whi = 0;
    if (L.n == 0 || dM < 0.0 || prec != prec) return;
    const float p1 = (float)((double)prec - dM);
    const float p2 = (float)((double)prec + dM);
    uint32_t lo, hi;
    dir_range(L.mdir, D, L.n, p1, &lo, &hi);
    uint32_t a = mass_lb(L.mass, lo, hi, p1);
    dir_range(L.mdir, D, L.n, p2, &lo, &hi);
    uint32_t b = mass_ub(L.mass, lo, hi, p2);
    *wlo = (int32_t)a;
    *whi = (int32_t)b - 1;
}

__device__ inline uint32_t iA_lower(const uint32_t *iA, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (iA[m] < key) lo = m + 1;
        else hi = m;
    }
    return lo;
}
__device__ inline uint32_t iA_upper(const uint32_t *iA, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (key < iA[m]) hi = m;
        else lo = m + 1;
    }
    return lo;
}

// minheap.h:108-177 (insert, increase_key at the root, heapify)
__device__ void heap_heapify(Cell *h, int size, int pos) {
    for (;;) {
        int r = (pos + 1) * 2, l = r - 1, s = pos;
        float sv = h[s].hs;
        if (l < size && h[l].hs < sv) { s = l; sv = h[l].hs; }
        if (r < size && h[r].hs < sv) { s = r; sv = h[r].hs; }
        if (s == pos) return;
        Cell t = h[s];
        h[s] = h[pos];
        h[pos] = t;
        pos = s;
    }
}
__device__ void heap_insert(Cell *h, int *size, int cap, const Cell &e) {
    if (*size == cap) {
        if (e.hs < h[0].hs) return;
        h[0] = e;
        heap_heapify(h, *size, 0);
        return;
    }
    h[(*size)++] = e;
    for (int i = (*size - 1) / 2; i >= 0; i--) heap_heapify(h, *size, i);
}
// minheap.h:258-277
__device__ Cell heap_getmax(const Cell *h, int size) {
    if (size <= 1) return h[0];
    Cell m = h[size - 1];
    for (int k = size - 2; k >= size / 2; k--)
        if (h[k].hs > m.hs) m = h[k];
    return m;
}

struct Smem {
    unsigned long long *sc;
    uint32_t *hist, *pcur, *pend, *pnxt, *pk_moz, *pk_int;
    int32_t *wlo, *whi;
    double *lgf;
    Cell *heap;
    float *rb_hs;
    uint32_t *rb_bc, *rb_yc;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline size_t smem_layout(const QueryArgs &a, bool replay, unsigned char *base, Smem *sm) {
    const uint32_t NP = a.nprobe_max;
    const uint32_t PM = NP / (2 * a.dF + 1);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align16(off + bytes);
        return base ? base + o : nullptr;
    };
    unsigned char *p;
    p = take(sizeof(unsigned long long) * 2 * a.tile);
    if (sm) sm->sc = (unsigned long long *)p;
    p = take(4ull * a.hbins);
    if (sm) sm->hist = (uint32_t *)p;
    p = take(4ull * NP);
    if (sm) sm->pcur = (uint32_t *)p;
    p = take(4ull * NP);
    if (sm) sm->pend = (uint32_t *)p;
    p = take(4ull * NP);
    if (sm) sm->pnxt = (uint32_t *)p;
    p = take(4ull * PM);
    if (sm) sm->pk_moz = (uint32_t *)p;
    p = take(4ull * PM);
    if (sm) sm->pk_int = (uint32_t *)p;
    p = take(4ull * a.nlen);
    if (sm) sm->wlo = (int32_t *)p;
    p = take(4ull * a.nlen);
    if (sm) sm->whi = (int32_t *)p;
    p = take(8ull * (a.max_shdpeaks + 1));
    if (sm) sm->lgf = (double *)p;
    if (replay) {
        p = take(sizeof(Cell) * (a.heap_cap + 1));
        if (sm) sm->heap = (Cell *)p;
        p = take(4ull * kBT);
        if (sm) sm->rb_hs = (float *)p;
        p = take(4ull * kBT);
        if (sm) sm->rb_bc = (uint32_t *)p;
        p = take(4ull * kBT);
        if (sm) sm->rb_yc = (uint32_t *)p;
    }
    return off;
}

template <bool REPLAY>
__global__ __launch_bounds__(kBT) void k_query(QueryArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ uint32_t s_maxbits, s_ties, s_cpsms, s_flags;
    __shared__ unsigned long long s_nia, s_nprobe;
    __shared__ Cell s_top;
    __shared__ int s_hsize;
    Smem sm;
    smem_layout(a, REPLAY, smem_raw, &sm);
    const int tid = threadIdx.x;
    const uint32_t W = 2 * a.dF + 1;
    const uint32_t D = a.max_mass + 2;
    for (uint32_t i = tid; i <= a.max_shdpeaks; i += kBT) sm.lgf[i] = a.lgf[i];
    for (uint32_t i = tid; i < 2 * a.tile; i += kBT) sm.sc[i] = 0;

    const uint32_t nwork = REPLAY ? a.tie[0] : (a.redo ? a.redo[0] : (uint32_t)a.nspec);
    for (uint32_t wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const int32_t s = REPLAY ? (int32_t)a.tie[1 + wi] : (int32_t)(a.redo ? a.redo[1 + wi] : wi);
        const uint32_t q0 = a.qidx[s];
        const uint32_t P = a.qidx[s + 1] - q0;
        const float prec = a.prec[s];
        for (uint32_t k = tid; k < P; k += kBT) {
            sm.pk_moz[k] = a.moz[q0 + k];
            sm.pk_int[k] = a.inten[q0 + k];
        }
        for (uint32_t h = tid; h < a.hbins; h += kBT) sm.hist[h] = 0;
        for (int l = tid; l < a.nlen; l += kBT) window(a.lens[l], prec, a.dM, D, &sm.wlo[l], &sm.whi[l]);
        if (tid == 0) {
            s_maxbits = 0;
            s_ties = 0;
            s_cpsms = 0;
            s_flags = 0;
            s_nia = 0;
            s_nprobe = 0;
            s_hsize = 0;
        }
        __syncthreads();

        float best_hs = 0.0f;
        uint32_t best_cnt = 0;
        Cell best = {0.0f, 0, 0, 0, 0, 0};
        uint32_t cpsms = 0, flags = 0;
        unsigned long long nia = 0, nprobe = 0;
        const uint32_t NP = P * W;

        for (int ixx = 0; ixx < a.nlen; ixx++) {
            const int32_t lo = sm.wlo[ixx], hi = sm.whi[ixx];
            if (lo > hi) continue;
            const LenDesc L = a.lens[ixx];
            const uint32_t speclen = L.speclen;
            const uint32_t lo_key = (uint32_t)lo * speclen;
            const uint32_t hi_key = (uint32_t)(hi + 1) * speclen - 1;
            for (uint32_t p = tid; p < NP; p += kBT) {
                const uint32_t k = p / W, j = p - k * W;
                const uint32_t q = sm.pk_moz[k];
                uint32_t c = 0, e = 0, nx = kNone;
                if (q > a.dF && q < a.NB - 1 - a.dF) {
                    const uint32_t bin = q - a.dF + j;
                    const uint32_t s0 = L.bA[bin], e0 = L.bA[bin + 1];
                    nprobe++;
                    if (e0 > s0) {
                        c = iA_lower(L.iA, s0, e0, lo_key);
                        e = iA_upper(L.iA, c, e0, hi_key);
                        nia += e - c;
                        if (c < e) nx = L.iA[c];
                    }
                }
                sm.pcur[p] = c;
                sm.pend[p] = e;
                sm.pnxt[p] = nx;
            }
            __syncthreads();
            for (int32_t t0 = lo; t0 <= hi; t0 += (int32_t)a.tile) {
                const uint32_t tn = (uint32_t)min((int32_t)a.tile, hi - t0 + 1);
                const uint32_t bound = (uint32_t)(t0 + (int32_t)tn) * speclen;
                for (uint32_t p = tid; p < NP; p += kBT) {
                    uint32_t nx = sm.pnxt[p];
                    if (nx >= bound) continue;
                    uint32_t c = sm.pcur[p];
                    const uint32_t e = sm.pend[p];
                    const unsigned long long add = kCnt + (unsigned long long)sm.pk_int[p / W];
                    do {
                        const uint32_t ppid = udiv(nx, L.div_m, L.div_s);
                        const uint32_t rem = nx - ppid * speclen;
                        const uint32_t isY = rem >= L.half ? 1u : 0u;
                        const uint32_t local = ppid - (uint32_t)t0;
                        atomicAdd(&sm.sc[2 * local + isY], add);
                        c++;
                        nx = (c < e) ? L.iA[c] : kNone;
                    } while (nx < bound);
                    sm.pcur[p] = c;
                    sm.pnxt[p] = nx;
                }
                __syncthreads();
                for (uint32_t cb = 0; cb < tn; cb += kBT) {
                    const uint32_t cc = cb + tid;
                    float hs_ins = -1.0f;
                    uint32_t bc = 0, yc = 0;
                    if (cc < tn) {
                        const unsigned long long vb = sm.sc[2 * cc], vy = sm.sc[2 * cc + 1];
                        if (vb | vy) {
                            sm.sc[2 * cc] = 0;
                            sm.sc[2 * cc + 1] = 0;
                            bc = (uint32_t)(vb >> 48);
                            yc = (uint32_t)(vy >> 48);
                            const uint16_t shpk = (uint16_t)(bc + yc);
                            if (shpk >= a.min_shp) {
                                uint32_t bi = bc, yi = yc;
                                if (bi > a.max_shdpeaks) { bi = a.max_shdpeaks; flags |= HCX_RES_LGF_CLAMP; }
                                if (yi > a.max_shdpeaks) { yi = a.max_shdpeaks; flags |= HCX_RES_LGF_CLAMP; }
                                double t = sm.lgf[bi] + sm.lgf[yi];
                                t = t + log10((double)(1ull + (vb & kSumMask)));
                                t = t + log10((double)(1ull + (vy & kSumMask)));
                                t = t - 6.0;
                                const float hs = (float)t;
                                if (hs > 0.0f) {
                                    cpsms++;
                                    int bin = (int)((double)(hs * 10.0f) + 0.5);
                                    if (bin >= (int)a.hbins) { bin = (int)a.hbins - 1; flags |= HCX_RES_HIST_CLAMP; }
                                    atomicAdd(&sm.hist[bin], 1u);
                                    hs_ins = hs;
                                    if (hs > best_hs) {
                                        best_hs = hs;
                                        best_cnt = 1;
                                        best = {hs, t0 + (int32_t)cc, (uint32_t)ixx, bc, yc, speclen};
                                    } else if (hs == best_hs) {
                                        best_cnt++;
                                    }
                                }
                            }
                        }
                    }
                    if (REPLAY) {
                        sm.rb_hs[tid] = hs_ins;
                        sm.rb_bc[tid] = bc;
                        sm.rb_yc[tid] = yc;
                        __syncthreads();
                        if (tid == 0) {
                            const uint32_t m = min((uint32_t)kBT, tn - cb);
                            for (uint32_t i = 0; i < m; i++) {
                                if (sm.rb_hs[i] > 0.0f) {
                                    Cell e = {sm.rb_hs[i], t0 + (int32_t)(cb + i), (uint32_t)ixx, sm.rb_bc[i],
                                              sm.rb_yc[i], speclen};
                                    heap_insert(sm.heap, &s_hsize, a.heap_cap, e);
                                }
                            }
                        }
                        __syncthreads();
                    }
                }
                __syncthreads();
            }
        }

        // ---- block reductions ----
        if (best_cnt) atomicMax(&s_maxbits, __float_as_uint(best_hs));
        if (cpsms) atomicAdd(&s_cpsms, cpsms);
        if (flags) atomicOr(&s_flags, flags);
        if (nia) atomicAdd(&s_nia, nia);
        if (nprobe) atomicAdd(&s_nprobe, nprobe);
        __syncthreads();
        const uint32_t maxb = s_maxbits;
        const bool mine = best_cnt && __float_as_uint(best_hs) == maxb;
        if (mine) atomicAdd(&s_ties, best_cnt);
        __syncthreads();
        if (!REPLAY && mine && s_ties == 1) s_top = best;
        if (REPLAY && tid == 0 && s_hsize > 0) s_top = heap_getmax(sm.heap, s_hsize);
        __syncthreads();
        if (tid == 0) {
            hcx_result r;
            r.hyperscore = 0.0f;
            r.psid = 0;
            r.idxoffset = 0;
            r.sharedions = 0;
            r.totalions = 0;
            r.bc = 0;
            r.yc = 0;
            r.flags = (uint16_t)s_flags;
            r.cpsms = s_cpsms;
            r.maxhypscore = r.mu = r.beta = r.minhypscore = r.nexthypscore = 0;
            r.e_x = 0.0;
            const bool have = REPLAY ? (s_hsize > 0) : (s_ties == 1);
            if (have) {
                r.hyperscore = s_top.hs;
                r.psid = s_top.psid;
                r.idxoffset = (uint16_t)s_top.ixx;
                r.sharedions = (uint16_t)(s_top.bc + s_top.yc);
                r.totalions = (uint16_t)s_top.speclen;
                r.bc = s_top.bc;
                r.yc = s_top.yc;
            }
            if (a.redo) r.flags |= HCX_RES_COUNT_OVF;   // re-scored here
            if (REPLAY) {
                r.flags |= HCX_RES_TIE_REPLAYED;
            } else if (s_ties > 1) {
                // tied maximum: the replay kernel resolves the heap order
                r.hyperscore = __uint_as_float(maxb);
                const uint32_t slot = atomicAdd(&a.tie[0], 1u);
                a.tie[1 + slot] = (uint32_t)s;
            }
            a.res[s] = r;
            if (a.stats) {
                a.stats[s].n_ia = s_nia;
                a.stats[s].n_probe = s_nprobe;
            }
        }
        if (!REPLAY) {
            uint32_t *hout = a.hist + (uint64_t)s * a.hbins;
            for (uint32_t h = tid; h < a.hbins; h += kBT) hout[h] = sm.hist[h];
        }
        __syncthreads();
    }
}

// Candidate dump (parity tests): same scan, one block per spectrum, every
// inserted candidate appended with its exact counters.
__global__ __launch_bounds__(kBT) void k_candidates(QueryArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Smem sm;
    smem_layout(a, false, smem_raw, &sm);
    const int tid = threadIdx.x;
    const uint32_t W = 2 * a.dF + 1;
    const uint32_t D = a.max_mass + 2;
    for (uint32_t i = tid; i <= a.max_shdpeaks; i += kBT) sm.lgf[i] = a.lgf[i];
    for (uint32_t i = tid; i < 2 * a.tile; i += kBT) sm.sc[i] = 0;
    const int32_t s = blockIdx.x;
    const uint32_t q0 = a.qidx[s];
    const uint32_t P = a.qidx[s + 1] - q0;
    for (uint32_t k = tid; k < P; k += kBT) {
        sm.pk_moz[k] = a.moz[q0 + k];
        sm.pk_int[k] = a.inten[q0 + k];
    }
    for (int l = tid; l < a.nlen; l += kBT) window(a.lens[l], a.prec[s], a.dM, D, &sm.wlo[l], &sm.whi[l]);
    __syncthreads();
    const uint32_t NP = P * W;
    for (int ixx = 0; ixx < a.nlen; ixx++) {
        const int32_t lo = sm.wlo[ixx], hi = sm.whi[ixx];
        if (lo > hi) continue;
        const LenDesc L = a.lens[ixx];
        const uint32_t speclen = L.speclen;
        for (uint32_t p = tid; p < NP; p += kBT) {
            const uint32_t k = p / W, j = p - k * W;
            const uint32_t q = sm.pk_moz[k];
            uint32_t c = 0, e = 0, nx = kNone;
            if (q > a.dF && q < a.NB - 1 - a.dF) {
                const uint32_t bin = q - a.dF + j;
                const uint32_t s0 = L.bA[bin], e0 = L.bA[bin + 1];
                if (e0 > s0) {
                    c = iA_lower(L.iA, s0, e0, (uint32_t)lo * speclen);
                    e = iA_upper(L.iA, c, e0, (uint32_t)(hi + 1) * speclen - 1);
                    if (c < e) nx = L.iA[c];
                }
            }
            sm.pcur[p] = c;
            sm.pend[p] = e;
            sm.pnxt[p] = nx;
        }
        __syncthreads();
        for (int32_t t0 = lo; t0 <= hi; t0 += (int32_t)a.tile) {
            const uint32_t tn = (uint32_t)min((int32_t)a.tile, hi - t0 + 1);
            const uint32_t bound = (uint32_t)(t0 + (int32_t)tn) * speclen;
            for (uint32_t p = tid; p < NP; p += kBT) {
                uint32_t nx = sm.pnxt[p];
                if (nx >= bound) continue;
                uint32_t c = sm.pcur[p];
                const uint32_t e = sm.pend[p];
                const unsigned long long add = kCnt + (unsigned long long)sm.pk_int[p / W];
                do {
                    const uint32_t ppid = udiv(nx, L.div_m, L.div_s);
                    const uint32_t isY = (nx - ppid * speclen) >= L.half ? 1u : 0u;
                    atomicAdd(&sm.sc[2 * (ppid - (uint32_t)t0) + isY], add);
                    c++;
                    nx = (c < e) ? L.iA[c] : kNone;
                } while (nx < bound);
                sm.pcur[p] = c;
                sm.pnxt[p] = nx;
            }
            __syncthreads();
            for (uint32_t cc = tid; cc < tn; cc += kBT) {
                const unsigned long long vb = sm.sc[2 * cc], vy = sm.sc[2 * cc + 1];
                if (!(vb | vy)) continue;
                sm.sc[2 * cc] = 0;
                sm.sc[2 * cc + 1] = 0;
                const uint32_t bc = (uint32_t)(vb >> 48), yc = (uint32_t)(vy >> 48);
                if ((uint16_t)(bc + yc) < a.min_shp) continue;
                const uint32_t bi = min(bc, a.max_shdpeaks), yi = min(yc, a.max_shdpeaks);
                double t = sm.lgf[bi] + sm.lgf[yi];
                t = t + log10((double)(1ull + (vb & kSumMask)));
                t = t + log10((double)(1ull + (vy & kSumMask)));
                t = t - 6.0;
                const float hs = (float)t;
                if (!(hs > 0.0f)) continue;
                const unsigned long long slot = atomicAdd(a.ncand, 1ull);
                if (slot < a.cand_cap) {
                    hcx_cand o;
                    o.spec = (uint32_t)s;
                    o.idxoffset = (uint16_t)ixx;
                    o.bc = (uint16_t)bc;
                    o.psid = t0 + (int32_t)cc;
                    o.yc = (uint16_t)yc;
                    o.pad_ = 0;
                    o.ibc = vb & kSumMask;
                    o.iyc = vy & kSumMask;
                    o.hyperscore = hs;
                    o.pad2_ = 0;
                    a.cand[slot] = o;
                }
            }
            __syncthreads();
        }
    }
}

// ---- expeRT::ModelTailFit (expeRT.cpp:572-718), one wavefront per spectrum ----
struct FitOut {
    int32_t mu, beta, stt1, end1;
    uint32_t flags;
    bool fragile;   // another correct log10 could give other ints: fit it again on the host
    int32_t lm, rm; // n > 3: first index at or below the 22 % threshold, last at or above the 87 % one
};

// v is within tol (+ its own rounding) of an integer, so truncating another
// value within tol of it could give a different int
__device__ __forceinline__ bool near_int(double v, double tol) {
    return fabs(v - rint(v)) <= tol + 4.0 * DBL_EPSILON * fabs(v);
}

// The fit's range: end1 = rargmax(yy, 0, hyp-1, 1.0), the last p <= hyp-1
// with yy[p] >= 1 (else hyp-1); stt1 = argmax(yy, 0, end1, 1.0), the first p
// in [0, end1] with yy[p] >= 1 (else 0).  Wave-uniform.
__device__ void fit_range(const uint32_t *H, int32_t hyp, int *stt1_out, int *end1_out) {
    const int lane = threadIdx.x & 63;
    int end1 = hyp - 1;
    for (int top = hyp - 1; top >= 0; top -= 64) {
        const int p = top - lane;
        const unsigned long long m = __ballot(p >= 0 && H[p] >= 1u);
        if (m) {
            end1 = top - (__ffsll((long long)m) - 1);
            break;
        }
    }
    int stt1 = 0;
    for (int b = 0; b <= end1; b += 64) {
        const int p = b + lane;
        const unsigned long long m = __ballot(p <= end1 && H[p] >= 1u);
        if (m) {
            stt1 = b + (__ffsll((long long)m) - 1);
            break;
        }
    }
    if (stt1 == end1) end1 += 1;
    *stt1_out = stt1;
    *end1_out = end1;
}

// The survival function over the window before its log10: sx[i] = 1 -
// (inclusive count through bin stt1 + i) / cpsms, values above 1 set to
// 0.999; returns the replacement for values <= 0, sx at rargmax(sx, 0, n-1,
// 1e-4).  Wave-uniform; sx complete on return.
__device__ double fit_survival(const uint32_t *H, int stt1, int n, uint32_t cpsms, double *sx) {
    const int lane = threadIdx.x & 63;
    const double vaa = (double)(int)cpsms;
    // inclusive prefix sums of integer counts: exact in any order
    double carry = 0.0;
    for (int b = 0; b < n; b += 64) {
        const int i = b + lane;
        double x = (i < n) ? (double)H[stt1 + i] : 0.0;
        for (int off = 1; off < 64; off <<= 1) {
            double y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (i < n) {
            double t = (carry + x) / vaa;
            t = t + (-1.0);
            sx[i] = -t;
        }
        carry += __shfl(x, 63, 64);
    }
    __syncwarp();
    for (int i = lane; i < n; i += 64)
        if (sx[i] > 1.0) sx[i] = 0.999;
    __syncwarp();
    int rep = n - 1;
    for (int top = n - 1; top >= 0; top -= 64) {
        const int p = top - lane;
        const unsigned long long m = __ballot(p >= 0 && sx[p] >= 1e-4);
        if (m) {
            rep = top - (__ffsll((long long)m) - 1);
            break;
        }
    }
    const double rv = sx[rep];
    __syncwarp();
    return rv;
}

// Lane 0: the least-squares line through the log10 values y[0, m) at bins
// stt1 + mark + i (expeRT.cpp:686-704), with the bound of how far slope and
// intercept move under another log10 within kFitYErr; fragile when the ints
// (values x 1e6 truncated) could then differ.
__device__ void line_fit(const double *y, int m, int x0, double &mu_t, double &beta_t, double &e_mu,
                         double &e_beta) {
    // e_mu, e_beta: first order through the sums, plus one rounding
    // difference per operation
    e_mu = 0.0;
    e_beta = kFitYErr;
    if (m == 1) {
        mu_t = 0.0;
        beta_t = y[0];
    } else {
        double xbar = 0.0, ybar = 0.0, top = 0.0, bot = 0.0, sdx = 0.0, sab = 0.0;
        for (int i = 0; i < m; i++) {
            xbar = xbar + (double)(x0 + i);
            ybar = ybar + y[i];
        }
        xbar = xbar / (double)m;
        ybar = ybar / (double)m;
        for (int i = 0; i < m; i++) {
            const double dx = (double)(x0 + i) - xbar;
            top = top + dx * (y[i] - ybar);
            bot = bot + dx * dx;
            sdx += fabs(dx);
            sab += fabs(dx * (y[i] - ybar));
        }
        mu_t = top / bot;
        beta_t = ybar - mu_t * xbar;
        e_mu = (2.0 * kFitYErr * sdx + (double)(m + 2) * DBL_EPSILON * sab) / bot + 2.0 * DBL_EPSILON * fabs(mu_t);
        e_beta = kFitYErr + (double)m * DBL_EPSILON * fabs(ybar) + e_mu * fabs(xbar) +
                 2.0 * DBL_EPSILON * (fabs(mu_t * xbar) + fabs(beta_t));
    }
}

// The one-rank fit's line: slope and intercept become ints x 1e6, so the
// fit is fragile when either lies within its bound of an integer there.
__device__ void fit_line(const double *y, int m, int x0, double &mu_t, double &beta_t, bool &fragile,
                         uint32_t &flags) {
    double e_mu, e_beta;
    line_fit(y, m, x0, mu_t, beta_t, e_mu, e_beta);
    // margin 4x; a single point's slope is an exact 0
    if ((m > 1 && near_int(mu_t * 1e6, 4e6 * e_mu)) || near_int(beta_t * 1e6, 4e6 * e_beta) ||
        fabs(fabs(mu_t) - 2.0e3) <= 4.0 * e_mu || fabs(fabs(beta_t) - 2.0e3) <= 4.0 * e_beta)
        fragile = true;
    if (!(fabs(mu_t) < 2.0e3) || !(fabs(beta_t) < 2.0e3)) {
        flags |= HCX_RES_FIT_UB;
        mu_t = 0.0;
        beta_t = 0.0;
    }
}

// sx: LDS scratch of this wave (>= hbins doubles).  Lane 0's return is valid.
__device__ FitOut tailfit_wave(const uint32_t *H, uint32_t hbins, uint32_t cpsms, int32_t hyp, double *sx) {
    const int lane = threadIdx.x & 63;
    FitOut o = {0, 0, 0, 0, 0, false, -1, -1};
    if (hyp < 1 || hyp >= (int)hbins) {
        // the reference reads yy[-1] or past the histogram (UB): a zero fit
        o.flags = HCX_RES_FIT_UB;
        return o;
    }
    int stt1, end1;
    fit_range(H, hyp, &stt1, &end1);
    const int n = end1 - stt1 + 1;
    o.stt1 = stt1;
    o.end1 = end1;
    double mu_t = 0.0, beta_t = 4.0;
    if (hyp < 1 || n < 2 || end1 >= (int)hbins) {
        o.flags |= HCX_RES_FIT_UB;
        mu_t = 0.0;
        beta_t = 0.0;
    } else if ((int)cpsms < 1) {
        mu_t = 0.0;
        beta_t = 100.0;
    } else {
        const double rv = fit_survival(H, stt1, n, cpsms, sx);
        for (int i = lane; i < n; i += 64) {
            double v = sx[i];
            if (v <= 0.0) v = rv;
            sx[i] = log10(v);
        }
        __syncwarp();
        const double s0 = sx[0];
        const double hgt = sx[n - 1] - s0;
        int mark = 0, mark2 = 0;
        auto largmax = [&](double v) {   // first i with sx[i] <= v, else 0
            for (int b = 0; b < n; b += 64) {
                const int i = b + lane;
                const unsigned long long m = __ballot(i < n && sx[i] <= v);
                if (m) return b + (__ffsll((long long)m) - 1);
            }
            return 0;
        };
        auto rargmax = [&](double v) {   // last i with sx[i] >= v, else n-1
            for (int top = n - 1; top >= 0; top -= 64) {
                const int p = top - lane;
                const unsigned long long m = __ballot(p >= 0 && sx[p] >= v);
                if (m) return top - (__ffsll((long long)m) - 1);
            }
            return n - 1;
        };
        // a log10 value this close to a window threshold could compare the
        // other way under another log10 within kFitYErr
        auto near_thr = [&](double v) {
            for (int b = 0; b < n; b += 64) {
                const int i = b + lane;
                if (__ballot(i < n && fabs(sx[i] - v) <= kFitWinEps)) return true;
            }
            return false;
        };
        bool fragile = false;
        if (n >= 3) fragile = near_thr(s0 + hgt * 0.22) || (n > 3 && near_thr(s0 + hgt * 0.87));
        if (n > 3) {
            o.lm = largmax(s0 + hgt * 0.22);
            o.rm = rargmax(s0 + hgt * 0.87);
            mark = o.lm - 1;
            mark2 = o.rm;
            if (mark2 == n) mark2 -= 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else if (n == 3) {
            mark = largmax(s0 + hgt * 0.22) - 1;
            mark2 = n - 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else {
            mark = 0;
            mark2 = n - 1;
        }
        if (mark < 0) {
            o.flags |= HCX_RES_FIT_UB;
            mu_t = 0.0;
            beta_t = 0.0;
        } else if (lane == 0) {
            fit_line(sx + mark, mark2 - mark + 1, stt1 + mark, mu_t, beta_t, fragile, o.flags);
        }
        o.fragile = fragile;
    }
    o.mu = (int32_t)(mu_t * 1e6);
    o.beta = (int32_t)(beta_t * 1e6);
    return o;
}

// an index whose log10 value is the constant tail's (z when the tail
// exists; otherwise any index, the value is then unused)
__device__ __forceinline__ int rep_of_tail(const double *, int z, int n) { return z < n ? z : n - 1; }

// The same fit with the window ends taken from a prediction (absolute bins
// p_lo ~ stt1 + lm, p_hi ~ stt1 + rm): the survival function is
// non-increasing up to its first value <= 0 (index z) and constant after
// (the replacement), so each end is the single threshold crossing of the
// monotone part, or, for the upper end, the constant tail itself.  The 64
// lanes test the 64 indices around each predicted end; a crossing found
// there is exactly the plain fit's, and the line is then fitted through the
// window alone.  Returns false (wave-uniform) when a crossing lies outside
// the tested indices or the histogram needs the plain fit's special cases
// (n <= 3, an undefined path); *o is then unset.
__device__ bool tailfit_window(const uint32_t *H, uint32_t hbins, uint32_t cpsms, int32_t hyp, double p_lo,
                               double p_hi, double *sx, FitOut *o) {
    const int lane = threadIdx.x & 63;
    if (hyp < 1 || hyp >= (int)hbins || (int)cpsms < 1 || !(p_lo == p_lo) || !(p_hi == p_hi)) return false;
    int stt1, end1;
    fit_range(H, hyp, &stt1, &end1);
    const int n = end1 - stt1 + 1;
    if (n <= 3 || end1 >= (int)hbins) return false;
    const double rv = fit_survival(H, stt1, n, cpsms, sx);
    int z = n;   // first value <= 0
    for (int b = 0; b < n; b += 64) {
        const int i = b + lane;
        const unsigned long long m = __ballot(i < n && sx[i] <= 0.0);
        if (m) {
            z = b + (__ffsll((long long)m) - 1);
            break;
        }
    }
    if (z == 0 || !(rv > 0.0)) return false;
    auto lg = [&](int i) {
        double v = sx[i];
        if (v <= 0.0) v = rv;
        return log10(v);
    };
    // lanes 0, 1, 2: the first and last values and the constant tail's
    // (indices >= z), in one log10 for the wave
    const double l3 = lg(lane == 0 ? 0 : lane == 1 ? n - 1 : rep_of_tail(sx, z, n));
    const double s0 = __shfl(l3, 0, 64);
    const double hgt = __shfl(l3, 1, 64) - s0;
    const double lr = __shfl(l3, 2, 64);
    const double th1 = s0 + hgt * 0.22, th2 = s0 + hgt * 0.87;
    if (!(th1 == th1) || !(th2 == th2) || !(lr == lr)) return false;
    // lower end: first index with value <= th1, a crossing inside [1, z);
    // lane l holds index c1 + l and tests it against lane l - 1's
    const double plo = fmin(fmax(rint(p_lo) - (double)stt1, 1.0), (double)(n - 1));
    const int c1 = (int)plo - 33;
    int lm = -1;
    bool frag = false;
    {
        const int q = c1 + lane;
        const double a = (q >= 0 && q < n) ? lg(q) : 0.0;
        const double b = __shfl_up(a, 1, 64);
        bool hit = false, near = false;
        if (lane >= 1 && q >= 1 && q < z) {
            hit = a <= th1 && b > th1;
            near = fabs(a - th1) <= kFitWinEps || fabs(b - th1) <= kFitWinEps;
        }
        const unsigned long long m = __ballot(hit);
        if (m) lm = c1 + (__ffsll((long long)m) - 1);
        frag = __ballot(hit && near) != 0;
    }
    if (lm < 0) return false;
    // upper end: last index with value >= th2: the end of the tail when the
    // tail's value reaches th2, else the crossing inside [0, z)
    int rm = -1;
    if (z < n) {
        if (fabs(lr - th2) <= kFitWinEps) frag = true;
        if (lr >= th2) rm = n - 1;
    }
    if (rm < 0) {
        // lane l holds index c2 + l and tests it against lane l + 1's
        const double phi = fmin(fmax(rint(p_hi) - (double)stt1, 0.0), (double)(n - 1));
        const int c2 = (int)phi - 31;
        const int q = c2 + lane;
        const double a = (q >= 0 && q < n) ? lg(q) : 0.0;
        const double b = __shfl_down(a, 1, 64);
        bool hit = false, near = false;
        if (q >= 0 && q < z) {
            if (q + 1 < z) {
                if (lane < 63) {
                    hit = a >= th2 && b < th2;
                    near = fabs(a - th2) <= kFitWinEps || fabs(b - th2) <= kFitWinEps;
                }
            } else {
                // q = z - 1: next is the tail (below th2 here) or the end
                hit = a >= th2;
                near = fabs(a - th2) <= kFitWinEps;
            }
        }
        const unsigned long long m = __ballot(hit);
        if (m) rm = c2 + (__ffsll((long long)m) - 1);
        if (__ballot(hit && near)) frag = true;
    }
    if (rm < 0) return false;
    FitOut r = {0, 0, stt1, end1, 0, frag, lm, rm};
    int mark = lm - 1, mark2 = rm;
    if (mark2 == n) mark2 -= 1;
    if (mark >= mark2) mark = mark2 - 1;
    double mu_t = 0.0, beta_t = 0.0;
    if (mark < 0) {
        r.flags |= HCX_RES_FIT_UB;
    } else {
        __syncwarp();
        for (int i = mark + lane; i <= mark2; i += 64) sx[i] = lg(i);   // each lane its own entries
        __syncwarp();
        if (lane == 0) fit_line(sx + mark, mark2 - mark + 1, stt1 + mark, mu_t, beta_t, r.fragile, r.flags);
    }
    r.mu = (int32_t)(mu_t * 1e6);
    r.beta = (int32_t)(beta_t * 1e6);
    *o = r;
    return true;
}

__device__ __forceinline__ void list_fragile(uint16_t &flags, uint32_t *fix, uint32_t s) {
    flags |= HCX_RES_FIT_HOST;
    if (fix) fix[1 + atomicAdd(fix, 1u)] = s;
}

__global__ __launch_bounds__(kBT) void k_tailfit(hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                                                 uint32_t min_cpsm, double expect_max, uint32_t *fix,
                                                 uint32_t force) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t s = blockIdx.x * (int32_t)(blockDim.x >> 6) + w;
    if (s >= n) return;
    double *sx = (double *)smem_raw + (size_t)w * hbins;
    hcx_result r = res[s];
    uint16_t flags = r.flags & ~(HCX_RES_VALID | HCX_RES_REPORTED | HCX_RES_SURROGATE);
    if (r.cpsms < min_cpsm || r.cpsms == 0) {
        if (lane == 0) res[s].flags = flags;
        return;
    }
    const int32_t hyp = (int32_t)((double)(r.hyperscore * 10.0f) + 0.5);
    FitOut f = tailfit_wave(hist + (size_t)s * hbins, hbins, r.cpsms, hyp, sx);
    if (lane == 0) {
        r.maxhypscore = hyp;
        r.mu = f.mu;
        r.beta = f.beta;
        r.minhypscore = f.stt1;
        r.nexthypscore = f.end1;
        double w_ = r.mu, b_ = r.beta;
        w_ /= 1e6;
        b_ /= 1e6;
        const double lgs = (w_ * r.maxhypscore) + b_;
        double e = pow(10.0, lgs);
        e *= (double)r.cpsms;
        r.e_x = e;
        flags |= (uint16_t)f.flags | HCX_RES_VALID;
        if (e < expect_max) flags |= HCX_RES_REPORTED;
        if (f.fragile || (force && s % force == 0)) list_fragile(flags, fix, (uint32_t)s);
        r.flags = flags;
        res[s] = r;
    }
}

// The plain fit of every stride-th spectrum with cpsms >= min_cpsm, compared
// with the surrogate e-value already in res[s]; the plain values replace it
// (flag HCX_RES_AUDITED).
__global__ __launch_bounds__(kBT) void k_tailfit_audit(hcx_result *res, const uint32_t *hist, uint32_t hbins,
                                                       int32_t n, uint32_t min_cpsm, double expect_max,
                                                       uint32_t stride, double tol, uint32_t *cnt, uint32_t *fix,
                                                       uint32_t force) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t s = ((int64_t)blockIdx.x * (blockDim.x >> 6) + w) * stride;
    if (s >= n) return;
    double *sx = (double *)smem_raw + (size_t)w * hbins;
    hcx_result r = res[s];
    if (r.cpsms < min_cpsm || r.cpsms == 0) return;
    const int32_t hyp = (int32_t)((double)(r.hyperscore * 10.0f) + 0.5);
    FitOut f = tailfit_wave(hist + (size_t)s * hbins, hbins, r.cpsms, hyp, sx);
    if (lane == 0) {
        double w_ = f.mu, b_ = f.beta;
        w_ /= 1e6;
        b_ /= 1e6;
        double e = pow(10.0, (w_ * hyp) + b_);
        e *= (double)r.cpsms;
        const double err = fabs(r.e_x - e) / fmax(fabs(e), 1e-300);
        atomicAdd(&cnt[0], 1u);
        if (!(err <= tol)) atomicAdd(&cnt[1], 1u);
        atomicMax((unsigned long long *)&cnt[2], (unsigned long long)__double_as_longlong(err == err ? err : 1e300));
        r.maxhypscore = hyp;
        r.mu = f.mu;
        r.beta = f.beta;
        r.minhypscore = f.stt1;
        r.nexthypscore = f.end1;
        r.e_x = e;
        uint16_t flags = r.flags & ~(HCX_RES_REPORTED | HCX_RES_SURROGATE);
        flags |= (uint16_t)f.flags | HCX_RES_VALID | HCX_RES_AUDITED;
        if (e < expect_max) flags |= HCX_RES_REPORTED;
        if (f.fragile || (force && s % force == 0)) list_fragile(flags, fix, (uint32_t)s);
        r.flags = flags;
        res[s] = r;
    }
}

__global__ __launch_bounds__(kBT) void k_tailfit_raw(const uint32_t *hist, uint32_t hbins, uint32_t n,
                                                     const uint32_t *cpsms, const int32_t *maxhyp, int32_t *out4,
                                                     uint32_t *flags, uint32_t force) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t s = blockIdx.x * (blockDim.x >> 6) + w;
    if (s >= n) return;
    double *sx = (double *)smem_raw + (size_t)w * hbins;
    FitOut f = tailfit_wave(hist + (size_t)s * hbins, hbins, cpsms[s], maxhyp[s], sx);
    if (lane == 0) {
        out4[4 * s + 0] = f.mu;
        out4[4 * s + 1] = f.beta;
        out4[4 * s + 2] = f.stt1;
        out4[4 * s + 3] = f.end1;
        if (flags) flags[s] = f.flags | ((f.fragile || (force && s % force == 0)) ? HCX_RES_FIT_HOST : 0u);
    }
}

// The plain fit's window ends (absolute bins stt1 + lm, stt1 + rm) of every
// spectrum with a regular fit (n > 3, defined path), -1 otherwise: the
// surrogate's training targets.  win: 2 per spectrum.
__global__ __launch_bounds__(kBT) void k_fit_windows(const hcx_result *res, const uint32_t *hist, uint32_t hbins,
                                                     int32_t n, uint32_t min_cpsm, int32_t *win) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t s = blockIdx.x * (int32_t)(blockDim.x >> 6) + w;
    if (s >= n) return;
    double *sx = (double *)smem_raw + (size_t)w * hbins;
    const hcx_result r = res[s];
    int32_t lo = -1, hi = -1;
    if (r.cpsms >= min_cpsm && r.cpsms > 0) {
        const int32_t hyp = (int32_t)((double)(r.hyperscore * 10.0f) + 0.5);
        const FitOut f = tailfit_wave(hist + (size_t)s * hbins, hbins, r.cpsms, hyp, sx);
        if (f.lm >= 1 && f.rm >= 0 && !(f.flags & HCX_RES_FIT_UB)) {
            lo = f.stt1 + f.lm;
            hi = f.stt1 + f.rm;
        }
    }
    if (lane == 0) {
        win[2 * s] = lo;
        win[2 * s + 1] = hi;
    }
}

// The surrogate's e-values: per spectrum, the predicted window ends (rows of
// Y, fractions of the top score's bin, standardised by norm) located
// exactly by tailfit_window, else the plain fit (tailfit_wave); either way
// the plain fit's ints.  cnt: per slot (workgroup modulo kSurSlots, slots
// kSurStride words apart) the spectra whose window the prediction located
// and those that took the plain fit.
__global__ __launch_bounds__(kBT) void k_sur_fit(const float *Y, uint32_t ystride, const float *norm,
                                                 hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                                                 uint32_t min_cpsm, double expect_max, uint32_t *fix, uint32_t force,
                                                 uint32_t *cnt) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t s = blockIdx.x * (int32_t)(blockDim.x >> 6) + w;
    if (s >= n) return;
    double *sx = (double *)smem_raw + (size_t)w * hbins;
    hcx_result r = res[s];
    uint16_t flags = r.flags & ~(HCX_RES_VALID | HCX_RES_REPORTED | HCX_RES_SURROGATE | HCX_RES_AUDITED);
    if (r.cpsms < min_cpsm || r.cpsms == 0) {
        if (lane == 0) res[s].flags = flags;
        return;
    }
    const int32_t hyp = (int32_t)((double)(r.hyperscore * 10.0f) + 0.5);
    const float *y = Y + (size_t)s * ystride;
    const double p_lo = ((double)y[0] * norm[3] + norm[0]) * hyp;
    const double p_hi = ((double)y[1] * norm[4] + norm[1]) * hyp;
    const uint32_t *H = hist + (size_t)s * hbins;
    FitOut f;
    const bool located = tailfit_window(H, hbins, r.cpsms, hyp, p_lo, p_hi, sx, &f);
    if (!located) {
        __syncwarp();
        f = tailfit_wave(H, hbins, r.cpsms, hyp, sx);
    }
    if (lane == 0) {
        atomicAdd(&cnt[(blockIdx.x % kSurSlots) * kSurStride + (located ? 0 : 1)], 1u);
        r.maxhypscore = hyp;
        r.mu = f.mu;
        r.beta = f.beta;
        r.minhypscore = f.stt1;
        r.nexthypscore = f.end1;
        double w_ = r.mu, b_ = r.beta;
        w_ /= 1e6;
        b_ /= 1e6;
        const double lgs = (w_ * r.maxhypscore) + b_;
        double e = pow(10.0, lgs);
        e *= (double)r.cpsms;
        r.e_x = e;
        flags |= (uint16_t)f.flags | HCX_RES_VALID | (located ? HCX_RES_SURROGATE : 0);
        if (e < expect_max) flags |= HCX_RES_REPORTED;
        if (f.fragile || (force && s % force == 0)) list_fragile(flags, fix, (uint32_t)s);
        r.flags = flags;
        res[s] = r;
    }
}

// ---- shard partial results and their merge (dslim_query.cpp:733-763,
// expeRT.cpp:720-936, dslim_score.cpp:244-345) ----

// (ushort) of a double as the reference's x86-64 code converts it: a 32-bit
// truncation, then the low 16 bits
__device__ __forceinline__ uint16_t u16_of(double v) { return (uint16_t)(uint32_t)(int32_t)v; }
// (int) of a double as x86-64's cvttsd2si converts it: out of range (or
// NaN) gives INT_MIN, where the device's conversion would saturate
__device__ __forceinline__ int32_t i32_of(double v) {
    return (v > -2147483649.0 && v < 2147483648.0) ? (int32_t)v : INT32_MIN;
}

// One wave per spectrum: the partRes and the u16 counts of the occupied
// histogram range (expeRT::StoreIResults), the rest of the slot zero.
__global__ __launch_bounds__(kBT) void k_part_export(const hcx_result *res, const uint32_t *hist, uint32_t hbins,
                                                     int32_t n, int32_t qid0, uint32_t slot, hcx_part_res *packs,
                                                     uint16_t *slices, uint32_t *overflow) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t s = blockIdx.x * (int32_t)(blockDim.x >> 6) + w;
    if (s >= n) return;
    const hcx_result r = res[s];
    const int32_t cpsms = (int32_t)r.cpsms;
    const uint32_t *yy = hist + (size_t)s * hbins;
    hcx_part_res pr;
    pr.min = 0;
    pr.max2 = 0;
    pr.max = 0.0f;
    pr.N = 0;
    pr.qID = qid0 + s;
    int stt = 0, ends = -1;
    if (cpsms >= 1) {
        // ends = rargmax(yy, 0, hbins-1, 0.99), stt = argmax(yy, 0, ends, 0.99)
        ends = (int)hbins - 1;
        for (int top = (int)hbins - 1; top >= 0; top -= 64) {
            const int p = top - lane;
            const unsigned long long m = __ballot(p >= 0 && yy[p] >= 1u);
            if (m) {
                ends = top - (__ffsll((long long)m) - 1);
                break;
            }
        }
        for (int b = 0; b <= ends; b += 64) {
            const int p = b + lane;
            const unsigned long long m = __ballot(p <= ends && yy[p] >= 1u);
            if (m) {
                stt = b + (__ffsll((long long)m) - 1);
                break;
            }
        }
        if ((uint32_t)(ends - stt + 1) > slot) {
            if (lane == 0) atomicOr(overflow, 1u);
            ends = -1;
        } else {
            pr.min = (uint16_t)stt;
            pr.max2 = (uint16_t)ends;
            pr.max = r.hyperscore;
            pr.N = cpsms;
        }
    }
    uint16_t *out = slices + (size_t)s * slot;
    for (uint32_t i = lane; i < slot; i += 64) {
        uint16_t k = 0;
        if ((int)i <= ends - stt) {
            k = (uint16_t)yy[stt + i];
            if (cpsms > 65500) k = u16_of((double)(int32_t)((uint32_t)k * 65500u) / (double)cpsms);
        }
        out[i] = k;
    }
    if (lane == 0) packs[s] = pr;
}

// expeRT::ModelTailFit(double&, int) on the merged histogram pd (LDS, this
// wave's hbins doubles, overwritten): returns the e-value (lane 0's is
// valid); *fragile when the window or the e-value's truncation x 1e6 or
// its comparison with expect_max lies within the device log10's error
// bound of a boundary.
__device__ double tailfit_merged_wave(double *pd, uint32_t hbins, int32_t pN, int32_t hyp, double expect_max,
                                      bool *fragile, uint32_t *flags) {
    const int lane = threadIdx.x & 63;
    *fragile = false;
    if (hyp < 1 || hyp >= (int)hbins) {
        *flags |= HCX_RES_FIT_UB;
        return INFINITY;
    }
    int end1 = hyp - 1;   // rargmax(pd, 0, hyp-1, 1.0)
    for (int top = hyp - 1; top >= 0; top -= 64) {
        const int p = top - lane;
        const unsigned long long m = __ballot(p >= 0 && pd[p] >= 1.0);
        if (m) {
            end1 = top - (__ffsll((long long)m) - 1);
            break;
        }
    }
    int stt1 = 0;         // argmax(pd, 0, end1, 1.0)
    for (int b = 0; b <= end1; b += 64) {
        const int p = b + lane;
        const unsigned long long m = __ballot(p <= end1 && pd[p] >= 1.0);
        if (m) {
            stt1 = b + (__ffsll((long long)m) - 1);
            break;
        }
    }
    const int n = end1 - stt1 + 1;
    const int vaa = pN;
    double mu_t = 0.0, beta_t = 4.0, e_mu = 0.0, e_beta = 0.0;
    bool frag = false;
    if (vaa < 1) {
        mu_t = 0.0;
        beta_t = 100.0;
    } else {
        double *sx = pd + stt1;
        if (lane == 0) {
            // std::accumulate(range, 1) into an int; then divide and
            // partial_sum in order (one lane: the order fixes the rounding)
            int acc = 1;
            for (int i = 0; i < n; i++) acc = (int)((double)acc + sx[i]);
            const int vaa2 = acc > vaa ? acc : vaa;
            double run = 0.0;
            for (int i = 0; i < n; i++) {
                const double q = sx[i] / (double)vaa2;
                run = (i == 0) ? q : run + q;
                const double t = run + (-1.0);
                sx[i] = -t;
            }
        }
        __syncwarp();
        for (int i = lane; i < n; i += 64)
            if (sx[i] > 1.0) sx[i] = 0.999;
        __syncwarp();
        int rep = n - 1;
        for (int top = n - 1; top >= 0; top -= 64) {
            const int p = top - lane;
            const unsigned long long m = __ballot(p >= 0 && sx[p] >= 1e-4);
            if (m) {
                rep = top - (__ffsll((long long)m) - 1);
                break;
            }
        }
        const double rv = sx[rep];
        __syncwarp();
        for (int i = lane; i < n; i += 64) {
            double v = sx[i];
            if (v <= 0.0) v = rv;
            sx[i] = log10(v);
        }
        __syncwarp();
        const double s0 = sx[0];
        const double hgt = sx[n - 1] - s0;
        auto largmax = [&](double v) {
            for (int b = 0; b < n; b += 64) {
                const int i = b + lane;
                const unsigned long long m = __ballot(i < n && sx[i] <= v);
                if (m) return b + (__ffsll((long long)m) - 1);
            }
            return 0;
        };
        auto rargmax = [&](double v) {
            for (int top = n - 1; top >= 0; top -= 64) {
                const int p = top - lane;
                const unsigned long long m = __ballot(p >= 0 && sx[p] >= v);
                if (m) return top - (__ffsll((long long)m) - 1);
            }
            return n - 1;
        };
        auto near_thr = [&](double v) {
            for (int b = 0; b < n; b += 64) {
                const int i = b + lane;
                if (__ballot(i < n && fabs(sx[i] - v) <= kFitWinEps)) return true;
            }
            return false;
        };
        if (n >= 3) frag = near_thr(s0 + hgt * 0.22) || (n > 3 && near_thr(s0 + hgt * 0.87));
        int mark = 0, mark2 = 0;
        if (n > 3) {
            mark = largmax(s0 + hgt * 0.22) - 1;
            mark2 = rargmax(s0 + hgt * 0.87);
            if (mark2 == n) mark2 -= 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else if (n == 3) {
            mark = largmax(s0 + hgt * 0.22) - 1;
            mark2 = n - 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else {
            mark = 0;
            mark2 = n - 1;
        }
        if (mark < 0) {
            *flags |= HCX_RES_FIT_UB;   // the reference reads sx[-1]: a zero line
            mu_t = 0.0;
            beta_t = 0.0;
        } else if (lane == 0) {
            line_fit(sx + mark, mark2 - mark + 1, stt1 + mark, mu_t, beta_t, e_mu, e_beta);
        }
    }
    const double e = pow(10, hyp * mu_t + beta_t) * vaa;
    if (lane == 0) {
        // relative bound on e under another log10 and pow (a few ulp each)
        const double eb = 2.302585092994046 * (fabs((double)hyp) * e_mu + e_beta) + 8.0 * DBL_EPSILON;
        if (near_int(e * 1e6, 4.0 * eb * e * 1e6) || fabs(e - expect_max) <= 4.0 * eb * e) frag = true;
    }
    *fragile = frag;
    return e;
}

// One wave per spectrum: CombineResults over G shards' partial results
// (packs [G][n], slices [G][n][slot]) -- histograms summed in shard order,
// the key shard of the top hyperscore, the merged tail fit; spectra whose
// fit is fragile are listed in fix (count, then ids) for the host.
__global__ __launch_bounds__(kBT) void k_combine(const hcx_part_res *packs, const uint16_t *slices, uint32_t G,
                                                 int32_t n, uint32_t slot, uint32_t hbins, uint32_t min_cpsm,
                                                 double expect_max, hcx_fresult *out, int32_t *keys, uint32_t *fix) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t s = blockIdx.x * (int32_t)(blockDim.x >> 6) + w;
    if (s >= n) return;
    double *pd = (double *)smem_raw + (size_t)w * hbins;
    for (uint32_t i = lane; i < hbins; i += 64) pd[i] = 0.0;
    __syncwarp();
    int32_t cpsms = 0, pN = 0, key = (int32_t)G;
    float maxhyp = -1.0f;
    for (uint32_t g = 0; g < G; g++) {
        const hcx_part_res r = packs[(size_t)g * n + s];
        if (r.min == 0 && r.max == 0.0f && r.N == 0 && r.max2 == 0) continue;
        cpsms += r.N;
        if (r.N >= 1) {
            pN += r.N;
            const uint16_t *buf = slices + ((size_t)g * n + s) * slot;
            for (int jj = r.min + lane; jj <= (int)r.max2; jj += 64) {
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
    __syncwarp();
    hcx_fresult f;
    f.eValue = i32_of(expect_max * 1e6);
    f.specID = -1;
    f.npsms = 0;
    int32_t k = (int32_t)G;
    if (key < (int32_t)G && cpsms >= (int32_t)min_cpsm) {
        const int32_t hyp = (int32_t)((double)(maxhyp * 10.0f) + 0.5);
        bool fragile = false;
        uint32_t fl = 0;
        const double e = tailfit_merged_wave(pd, hbins, pN, hyp, expect_max, &fragile, &fl);
        if (lane == 0) {
            if (e < expect_max) {
                f.eValue = i32_of(e * 1e6);
                f.specID = packs[(size_t)key * n + s].qID;
                f.npsms = cpsms;
                k = key;
            }
            if (fragile && fix) fix[1 + atomicAdd(fix, 1u)] = (uint32_t)s;
        }
    }
    if (lane == 0) {
        out[s] = f;
        keys[s] = k;
    }
}

}  // namespace

size_t query_smem_bytes(const QueryArgs &a, bool replay) { return smem_layout(a, replay, nullptr, nullptr); }

// Cursor form: one workgroup per spectrum over the reference bA/iA only
// (per-probe binary searches, then cursors across LDS tiles).  Records
// ctx->ev0 / ev1 around the kernel.
int launch_cursor(hcx_ctx *ctx, const QueryArgs &a) {
    const size_t sm = query_smem_bytes(a, false);
    HCX_HIP(hipFuncSetAttribute((const void *)k_query<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    if (!ctx->capturing) HCX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (a.nspec > 0) {
        k_query<false><<<a.nspec, kBT, sm, ctx->stream>>>(a);
        HCX_HIP(hipGetLastError());
    }
    if (!ctx->capturing) HCX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    return HCX_SUCCESS;
}

// The cursor kernel over the spectra of a.redo, whose count only the device
// knows: a persistent grid.
int launch_cursor_redo(hcx_ctx *ctx, const QueryArgs &a) {
    const size_t sm = query_smem_bytes(a, false);
    HCX_HIP(hipFuncSetAttribute((const void *)k_query<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    const int grid = std::max(1, std::min(a.nspec, ctx->num_cu * 2));
    if (a.nspec > 0) {
        k_query<false><<<grid, kBT, sm, ctx->stream>>>(a);
        HCX_HIP(hipGetLastError());
    }
    return HCX_SUCCESS;
}

// Reference-order replay of tied spectra on the cursor path (kept as an
// independent second implementation for the parity tests).
int launch_replay_cursor(hcx_ctx *ctx, const QueryArgs &a) {
    const size_t sm = query_smem_bytes(a, true);
    HCX_HIP(hipFuncSetAttribute((const void *)k_query<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    const int grid = std::max(1, std::min(a.nspec, ctx->num_cu * 2));
    k_query<true><<<grid, kBT, sm, ctx->stream>>>(a);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_candidates(hcx_ctx *ctx, const QueryArgs &a) {
    const size_t sm = query_smem_bytes(a, false);
    HCX_HIP(hipFuncSetAttribute((const void *)k_candidates, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    if (a.nspec > 0) {
        k_candidates<<<a.nspec, kBT, sm, ctx->stream>>>(a);
        HCX_HIP(hipGetLastError());
    }
    return HCX_SUCCESS;
}

// Wavefronts per workgroup of the fit kernels: each holds hbins doubles of
// LDS, up to four per workgroup within the device's limit.
static int fit_waves(const hcx_ctx *ctx, uint32_t hbins) {
    const size_t per_wave = (size_t)hbins * sizeof(double);
    const int w = (int)std::min<size_t>(kBT / 64, (size_t)ctx->max_lds / per_wave);
    return std::max(1, w);
}

int launch_tailfit(hcx_ctx *ctx, hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                   uint32_t min_cpsm, double expect_max, uint32_t *fix, uint32_t force) {
    if (n <= 0) return HCX_SUCCESS;
    const int per = fit_waves(ctx, hbins);
    const size_t sm = (size_t)per * hbins * sizeof(double);
    HCX_HIP(hipFuncSetAttribute((const void *)k_tailfit, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_tailfit<<<(n + per - 1) / per, per * 64, sm, ctx->stream>>>(res, hist, hbins, n, min_cpsm, expect_max, fix,
                                                             force);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_tailfit_audit(hcx_ctx *ctx, hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                         uint32_t min_cpsm, double expect_max, uint32_t stride, double tol, uint32_t *cnt,
                         uint32_t *fix, uint32_t force) {
    if (n <= 0) return HCX_SUCCESS;
    const int per = fit_waves(ctx, hbins);
    const size_t sm = (size_t)per * hbins * sizeof(double);
    HCX_HIP(hipFuncSetAttribute((const void *)k_tailfit_audit, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    const int64_t m = ((int64_t)n + stride - 1) / stride;
    k_tailfit_audit<<<(unsigned)((m + per - 1) / per), per * 64, sm, ctx->stream>>>(res, hist, hbins, n, min_cpsm,
                                                                             expect_max, stride, tol, cnt, fix, force);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_tailfit_raw(hcx_ctx *ctx, const uint32_t *hist, uint32_t hbins, uint32_t n, const uint32_t *cpsms,
                       const int32_t *maxhyp, int32_t *out4, uint32_t *flags, uint32_t force) {
    if (n == 0) return HCX_SUCCESS;
    const uint32_t per = (uint32_t)fit_waves(ctx, hbins);
    const size_t sm = (size_t)per * hbins * sizeof(double);
    HCX_HIP(hipFuncSetAttribute((const void *)k_tailfit_raw, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_tailfit_raw<<<(n + per - 1) / per, per * 64, sm, ctx->stream>>>(hist, hbins, n, cpsms, maxhyp, out4, flags,
                                                                 force);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_fit_windows(hcx_ctx *ctx, const hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                       uint32_t min_cpsm, int32_t *win) {
    if (n <= 0) return HCX_SUCCESS;
    const int per = fit_waves(ctx, hbins);
    const size_t sm = (size_t)per * hbins * sizeof(double);
    HCX_HIP(hipFuncSetAttribute((const void *)k_fit_windows, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_fit_windows<<<(n + per - 1) / per, per * 64, sm, ctx->stream>>>(res, hist, hbins, n, min_cpsm, win);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_sur_fit(hcx_ctx *ctx, const float *Y, uint32_t ystride, const float *norm, hcx_result *res,
                   const uint32_t *hist, uint32_t hbins, int32_t n, uint32_t min_cpsm, double expect_max,
                   uint32_t *fix, uint32_t force, uint32_t *cnt) {
    if (n <= 0) return HCX_SUCCESS;
    const int per = fit_waves(ctx, hbins);
    const size_t sm = (size_t)per * hbins * sizeof(double);
    HCX_HIP(hipFuncSetAttribute((const void *)k_sur_fit, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_sur_fit<<<(n + per - 1) / per, per * 64, sm, ctx->stream>>>(Y, ystride, norm, res, hist, hbins, n, min_cpsm,
                                                                  expect_max, fix, force, cnt);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_part_export(hcx_ctx *ctx, const hcx_result *res, const uint32_t *hist, uint32_t hbins, int32_t n,
                       int32_t qid0, uint32_t slot, hcx_part_res *packs, uint16_t *slices, uint32_t *overflow) {
    if (n <= 0) return HCX_SUCCESS;
    const int per = kBT / 64;
    k_part_export<<<(n + per - 1) / per, kBT, 0, ctx->stream>>>(res, hist, hbins, n, qid0, slot, packs, slices,
                                                                overflow);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_combine(hcx_ctx *ctx, const hcx_part_res *packs, const uint16_t *slices, uint32_t G, int32_t n,
                   uint32_t slot, hcx_fresult *out, int32_t *keys, uint32_t *fix) {
    if (n <= 0) return HCX_SUCCESS;
    const int per = fit_waves(ctx, ctx->hbins);
    const size_t sm = (size_t)per * ctx->hbins * sizeof(double);
    HCX_HIP(hipFuncSetAttribute((const void *)k_combine, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_combine<<<(n + per - 1) / per, per * 64, sm, ctx->stream>>>(packs, slices, G, n, slot, ctx->hbins,
                                                                  ctx->params.min_cpsm, ctx->params.expect_max, out,
                                                                  keys, fix);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

// Histogram rows and results of the listed spectra into staging (one
// workgroup per spectrum), and the corrected results back.
__global__ __launch_bounds__(256) void k_fix_gather(const uint32_t *ids, const uint32_t *hist, uint32_t hbins,
                                                    const hcx_result *res, uint32_t *rows, hcx_result *rstage) {
    const uint32_t s = ids[blockIdx.x];
    const uint32_t *src = hist + (size_t)s * hbins;
    uint32_t *dst = rows + (size_t)blockIdx.x * hbins;
    for (uint32_t i = threadIdx.x; i < hbins; i += 256) dst[i] = src[i];
    if (threadIdx.x == 0) rstage[blockIdx.x] = res[s];
}

__global__ __launch_bounds__(256) void k_fix_scatter(const uint32_t *ids, uint32_t k, const hcx_result *rstage,
                                                     hcx_result *res) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < k) res[ids[i]] = rstage[i];
}

int launch_fix_gather(hcx_ctx *c, const uint32_t *ids, uint32_t k, const uint32_t *hist, uint32_t hbins,
                      const hcx_result *res, uint32_t *rows, hcx_result *rstage) {
    if (k == 0) return HCX_SUCCESS;
    k_fix_gather<<<k, 256, 0, c->stream>>>(ids, hist, hbins, res, rows, rstage);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int launch_fix_scatter(hcx_ctx *c, const uint32_t *ids, uint32_t k, const hcx_result *rstage, hcx_result *res) {
    if (k == 0) return HCX_SUCCESS;
    k_fix_scatter<<<(k + 255) / 256, 256, 0, c->stream>>>(ids, k, rstage, res);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

}  // namespace hcx
