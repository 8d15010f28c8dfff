// k_index.hip -- DSLIM index build on the GPU (DSLIM_Construct, dslim.cpp:46-138).
//
// 1. k_frag: one thread per (peptide form, charge, series) walks the L-1
//    fragment bins of that series (UTILS_GenerateSpectrum utils.cpp:231-276,
//    UTILS_GenerateModSpectrum utils.cpp:378-505; dslim.cpp:300-323).  FP64
//    with contraction off, so the truncated bins equal the host's.
// 2. iA = ion ids stably sorted by bin: LSD radix sort, 8-bit digits, tiles of
//    4096 keys, wave multi-split ranking (ballots) so equal digits keep their
//    input order.  A stable sort of ascending ids equals the reference's
//    quicksort followed by the per-bin ascending sort (dslim.cpp:361-388,
//    452-482).
// 3. bA[b] = ions in bins below b, the exclusive scan of the bin histogram
//    (dslim.cpp:89-126), read off the sorted bins with no atomics.
// 4. mdir[d] = first entry with mass >= d Da: a directory that shortens the
//    precursor-window binary searches of the query.
#include "hcx_internal.h"
#include "k_mass.h"

namespace hcx {
namespace {

constexpr int kBT = 256;            // threads per block
constexpr int kIPT = 16;            // items per thread
constexpr int kTile = kBT * kIPT;   // 4096
constexpr int kRB = 8;              // radix bits per pass
constexpr int kR = 1 << kRB;

// One thread per (form, series, charge) row of L-1 bins; row t of the output
// starts at t * (L-1), so a block's rows are one contiguous range: the rows go
// to LDS (odd row stride, no bank conflicts) and out with coalesced stores.
__device__ __forceinline__ void frag_row(const BuildArgs &a, uint64_t t, uint32_t *out);

__global__ __launch_bounds__(kBT) void k_frag(BuildArgs a, uint32_t *keys) {
    extern __shared__ uint32_t stage[];
    const uint32_t per = a.maxz * 2;
    const uint32_t L1 = a.peplen - 1;
    const uint32_t S = L1 | 1u;
    const uint64_t rows = (uint64_t)a.n * per;
    const uint64_t t0 = (uint64_t)blockIdx.x * kBT;
    const uint64_t t = t0 + threadIdx.x;
    if (t < rows) frag_row(a, t, stage + threadIdx.x * S);
    __syncthreads();
    const uint32_t nr = (uint32_t)min<uint64_t>(kBT, rows - t0);
    uint32_t *dst = keys + t0 * L1;
    for (uint32_t j = threadIdx.x; j < nr * L1; j += kBT) {
        const uint32_t r = j / L1;
        dst[j] = stage[r * S + (j - r * L1)];
    }
}

__device__ __forceinline__ void frag_row(const BuildArgs &a, uint64_t t, uint32_t *out) {
    const uint32_t per = a.maxz * 2;
    const uint32_t k = (uint32_t)(t / per);
    const uint32_t r = (uint32_t)(t % per);
    const uint32_t series = r / a.maxz;      // 0 = b, 1 = y
    const uint32_t z = r % a.maxz;
    const uint32_t c = z + 1;
    const uint32_t L = a.peplen;
    const uint32_t L1 = L - 1;
    const hcx_pep_entry e = a.entries[k];
    const char *seq = a.seqs + (size_t)e.seq_id * L;

    const bool mod = e.mod_num != 0;
    float mass;
    if (!mod) {
        mass = pep_mass_dev(seq, L);
    } else if (e.sites == 0) {
        mass = -20000.0f;
    } else {
        mass = pep_mass_dev(seq, L);
        uint32_t first = e.mod_num & 0xF;     // first mod only (utils.cpp:350-358)
        if (first) mass += (float)a.mod_mass[first - 1] / (float)a.scale;
    }
    const bool legal = (double)mass >= a.min_mass && (double)mass <= a.max_mass;
    const bool filled = mod ? ((double)mass > a.min_mass && (double)mass < a.max_mass) : (mass > 0.0f);
    if (!legal || !filled) {
        // illegal form: speclen zeros into bin 0 (dslim.cpp:314-323); a legal
        // mod form the reference leaves unfilled is zero-filled too
        for (uint32_t l = 0; l < L1; l++) out[l] = 0;
        return;
    }
    const double scale = (double)a.scale;
    const uint32_t NB = a.NB;
    if (!mod) {
        uint32_t v;
        if (series == 0) {
            v = (uint32_t)((getaa(seq[0], c) * scale) / (double)c);
        } else {
            v = (uint32_t)(((getaa(seq[L - 1], c) + (double)18.015f) * scale) / (double)c);
        }
        for (uint32_t l = 0; l < L1; l++) {
            if (l > 0) {
                char ch = series == 0 ? seq[l] : seq[L - 1 - l];
                v += (uint32_t)((getaa(ch, 0) * scale) / (double)c);
            }
            uint32_t b = v >= NB ? NB - 1 : v;
            out[l] = b;
        }
    } else {
        // cumulative undivided sums + prefix of mod masses, then integer divide
        int mods[8];
        int seen = 0;
        for (int i = 0; i < 8; i++) {
            int nib = (int)((e.mod_num >> (4 * i)) & 0xF) - 1;
            mods[i] = nib;
            if (nib != -1) seen++;
        }
        uint32_t cum;
        if (series == 0) cum = (uint32_t)(getaa(seq[0], c) * scale);
        else cum = (uint32_t)((getaa(seq[L - 1], c) + (double)18.015f) * scale);
        uint32_t counter = 0, madd = 0;
        for (uint32_t l = 0; l < L1; l++) {
            uint32_t pos = series == 0 ? l : L - 1 - l;
            if (l > 0) cum += (uint32_t)(getaa(seq[pos], 0) * scale);
            if ((e.sites >> pos) & 1ull) {
                int mi = series == 0 ? mods[counter] : mods[seen - 1 - (int)counter];
                counter++;
                madd += a.mod_mass[mi];
            }
            uint32_t v = (cum + madd) / c;
            uint32_t b = v >= NB ? NB - 1 : v;
            out[l] = b;
        }
    }
}

// ---- exclusive scan of u32 (tiles of 4096; padded LDS, no bank conflicts) ----
__device__ inline uint32_t pad_idx(uint32_t i) { return i + (i >> 4); }

__global__ __launch_bounds__(kBT) void k_scan_reduce(const uint32_t *in, uint64_t n, uint32_t *sums) {
    __shared__ uint32_t red[kBT / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    uint32_t s = 0;
    for (int r = 0; r < kIPT; r++) {
        uint64_t i = base + (uint64_t)r * kBT + threadIdx.x;
        if (i < n) s += in[i];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBT / 64; w++) t += red[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBT) void k_scan_block(const uint32_t *in, uint32_t *out, uint64_t n,
                                                    const uint32_t *offs) {
    __shared__ uint32_t s[kTile + kTile / 16];
    __shared__ uint32_t wsum[kBT / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    for (int r = 0; r < kIPT; r++) {
        uint32_t li = r * kBT + threadIdx.x;
        uint64_t i = base + li;
        s[pad_idx(li)] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t loc[kIPT];
    uint32_t tsum = 0;
    for (int j = 0; j < kIPT; j++) {
        loc[j] = s[pad_idx(threadIdx.x * kIPT + j)];
        tsum += loc[j];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = tsum;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t run = offs ? offs[blockIdx.x] : 0u;
    for (int ww = 0; ww < w; ww++) run += wsum[ww];
    run += x - tsum;
    for (int j = 0; j < kIPT; j++) {
        s[pad_idx(threadIdx.x * kIPT + j)] = run;
        run += loc[j];
    }
    __syncthreads();
    for (int r = 0; r < kIPT; r++) {
        uint32_t li = r * kBT + threadIdx.x;
        uint64_t i = base + li;
        if (i < n) out[i] = s[pad_idx(li)];
    }
}

// ---- radix sort ----
__global__ __launch_bounds__(kBT) void k_radix_hist(const uint32_t *keys, uint64_t n, int shift,
                                                    uint32_t *tc, uint32_t ntiles) {
    __shared__ uint32_t h[kR];
    for (int d = threadIdx.x; d < kR; d += kBT) h[d] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    for (int r = 0; r < kIPT; r++) {
        uint64_t i = base + (uint64_t)r * kBT + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kR - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kR; d += kBT) tc[(uint64_t)d * ntiles + blockIdx.x] = h[d];
}

// One tile of kTile keys: stable ranks within the tile by digit (wave
// multi-split per round of kBT keys, rounds and waves in input order), the
// (key, value) pairs staged in LDS in digit order, then written out so that
// each digit's run of the tile goes to consecutive addresses.
static_assert(kBT == kR, "one thread per digit");

__global__ __launch_bounds__(kBT) void k_radix_scatter(const uint32_t *kin, const uint32_t *vin,
                                                       uint32_t *kout, uint32_t *vout, uint64_t n,
                                                       int shift, const uint32_t *offs, uint32_t ntiles) {
    __shared__ uint32_t start[kR];        // tile-local start of each digit's run
    __shared__ uint32_t run[kR];          // next tile-local slot of each digit
    __shared__ uint32_t gofs[kR];         // global start of (digit, tile)
    __shared__ uint32_t wc[kBT / 64][kR];
    __shared__ uint32_t sk[kTile], sv[kTile];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t tbase = (uint64_t)blockIdx.x * kTile;
    const uint32_t tn = (uint32_t)min<uint64_t>(kTile, n - tbase);
    {
        // the tile's count of digit d, from the scanned (digit, tile) offsets
        // (digit-major: the next entry is the next tile's or the next digit's)
        const int d = threadIdx.x;
        const uint64_t at = (uint64_t)d * ntiles + blockIdx.x;
        const uint32_t g = offs[at];
        const uint32_t cnt = (at + 1 < (uint64_t)kR * ntiles ? offs[at + 1] : (uint32_t)n) - g;
        gofs[d] = g;
        run[d] = cnt;
        for (int ww = 0; ww < kBT / 64; ww++) wc[ww][d] = 0;
        __syncthreads();
        // exclusive scan of the counts over the digits
        for (int off = 1; off < kR; off <<= 1) {
            const uint32_t v = d >= off ? run[d - off] : 0u;
            __syncthreads();
            run[d] += v;
            __syncthreads();
        }
        start[d] = run[d] - cnt;
        run[d] = start[d];
        __syncthreads();
    }
    for (int r = 0; r < kIPT; r++) {
        const uint64_t i = tbase + (uint64_t)r * kBT + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t val = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
        const uint32_t d = (key >> shift) & (kR - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRB; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int ww = 0; ww < w; ww++) pos += wc[ww][d];
            sk[pos] = key;
            sv[pos] = val;
        }
        __syncthreads();
        {
            const int dd = threadIdx.x;
            uint32_t add = 0;
            for (int ww = 0; ww < kBT / 64; ww++) {
                add += wc[ww][dd];
                wc[ww][dd] = 0;
            }
            run[dd] += add;
        }
        __syncthreads();
    }
    // staged entry j of digit d goes to gofs[d] + (j - start[d]): each
    // digit's run of the tile lands on consecutive addresses
    for (uint32_t j = threadIdx.x; j < tn; j += kBT) {
        const uint32_t key = sk[j];
        const uint32_t d = (key >> shift) & (kR - 1);
        const uint32_t pos = gofs[d] + (j - start[d]);
        if (kout) kout[pos] = key;
        vout[pos] = sv[j];
    }
}

// bA[b] = number of ions in bins below b (dslim.cpp:89-126): the lower bound
// of b in the bins sorted ascending, one thread per bin (neighbouring bins
// take nearly the same search path).
__global__ void k_bucket_offsets(const uint32_t *skeys, uint64_t n, uint32_t NB, uint32_t *bA) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > NB) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t m = (lo + hi) >> 1;
        if (skeys[m] < b) lo = m + 1;
        else hi = m;
    }
    bA[b] = (uint32_t)lo;
}

__global__ void k_mdir(const float *mass, uint32_t n, uint32_t *mdir, uint32_t D) {
    uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    const float v = (float)d;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (mass[m] < v) lo = m + 1;
        else hi = m;
    }
    mdir[d] = lo;
}

}  // namespace

// Exclusive scan of n u32 on the stream; tmp must hold scan_tmp_words(n) u32.
uint64_t scan_tmp_words(uint64_t n) {
    uint64_t w = 0;
    while (n > (uint64_t)kTile) {
        n = (n + kTile - 1) / kTile;
        w += n + 1;
    }
    return w + 1;
}

int scan_exclusive_u32(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tmp);
static int scan_exclusive(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tmp) {
    return scan_exclusive_u32(st, in, out, n, tmp);
}
int scan_exclusive_u32(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tmp) {
    if (n == 0) return HCX_SUCCESS;
    const uint64_t nb = (n + kTile - 1) / kTile;
    if (nb == 1) {
        k_scan_block<<<1, kBT, 0, st>>>(in, out, n, nullptr);
        HCX_HIP(hipGetLastError());
        return HCX_SUCCESS;
    }
    uint32_t *sums = tmp;
    k_scan_reduce<<<(unsigned)nb, kBT, 0, st>>>(in, n, sums);
    HCX_HIP(hipGetLastError());
    int rc = scan_exclusive_u32(st, sums, sums, nb, tmp + nb + 1);
    if (rc) return rc;
    k_scan_block<<<(unsigned)nb, kBT, 0, st>>>(in, out, n, sums);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

namespace {

struct BuildSizes {
    int passes;
    uint64_t ntiles, tc_words, scan_words;
};

BuildSizes build_sizes(uint64_t nions, uint32_t NB) {
    uint32_t keybits = 1;
    while (keybits < 32 && (1ull << keybits) < NB) keybits++;
    BuildSizes z;
    z.passes = (int)((keybits + kRB - 1) / kRB);
    z.ntiles = (nions + kTile - 1) / kTile;
    z.tc_words = (uint64_t)kR * (z.ntiles ? z.ntiles : 1);
    z.scan_words = scan_tmp_words(z.tc_words);
    return z;
}

// grow-only, with headroom, so that a run of lengths of rising size
// reallocates only a few times
int grow(DeviceBuf &d, uint64_t bytes) {
    if (bytes <= d.bytes && d.p) return 0;
    return d.ensure(std::max<uint64_t>(bytes, d.bytes + d.bytes / 2));
}

}  // namespace

int reserve_index_build(hcx_ctx *ctx, uint64_t nions, uint32_t NB) {
    const BuildSizes z = build_sizes(nions, NB);
    const bool fits = ctx->bld[0].bytes >= nions * 4 + 4 && ctx->bld[4].bytes >= z.tc_words * 4 &&
                      ctx->bld[5].bytes >= z.scan_words * 4 && (z.passes <= 2 || ctx->bld[3].bytes >= nions * 4 + 4);
    if (fits) return HCX_SUCCESS;
    // nothing queued may still use a buffer that is about to be replaced
    HCX_HIP(hipStreamSynchronize(ctx->stream));
    if (grow(ctx->bld[0], nions * 4 + 4) || grow(ctx->bld[1], nions * 4 + 4) || grow(ctx->bld[2], nions * 4 + 4) ||
        (z.passes > 2 && grow(ctx->bld[3], nions * 4 + 4)) || grow(ctx->bld[4], z.tc_words * 4) ||
        grow(ctx->bld[5], z.scan_words * 4))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "index build: device allocation failed");
    return HCX_SUCCESS;
}

int launch_index_build(hcx_ctx *ctx, const BuildArgs &a, uint32_t *d_iA, uint32_t *d_bA, const float *d_mass,
                       uint32_t *d_mdir, uint32_t max_mass) {
    hipStream_t st = ctx->stream;
    const uint64_t nions = (uint64_t)a.n * a.speclen;
    const uint32_t NB = a.NB;
    const BuildSizes z = build_sizes(nions, NB);
    const int passes = z.passes;
    const uint64_t ntiles = z.ntiles, tc_words = z.tc_words;
    int rc0 = reserve_index_build(ctx, nions, NB);
    if (rc0) return rc0;
    DeviceBuf &K0 = ctx->bld[0], &K1 = ctx->bld[1], &V1 = ctx->bld[2], &V2 = ctx->bld[3], &tc = ctx->bld[4],
              &tmp = ctx->bld[5];
    uint32_t *k0 = (uint32_t *)K0.p, *k1 = (uint32_t *)K1.p, *v1 = (uint32_t *)V1.p, *v2 = (uint32_t *)V2.p;
    uint32_t *dtc = (uint32_t *)tc.p, *dtmp = (uint32_t *)tmp.p;

    const uint64_t nthr = (uint64_t)a.n * a.maxz * 2;
    if (nthr) {
        const size_t lds = (size_t)kBT * ((a.peplen - 1) | 1u) * 4;
        k_frag<<<(unsigned)((nthr + kBT - 1) / kBT), kBT, lds, st>>>(a, k0);
        HCX_HIP(hipGetLastError());
    }
    int rc = HCX_SUCCESS;
    const uint32_t *sorted = k0;   // the bins in ascending order once sorted
    if (nions) {
        const uint32_t *kin = k0, *vin = nullptr;
        for (int p = 0; p < passes; p++) {
            const bool last = p == passes - 1;
            uint32_t *kout = kin == k0 ? k1 : k0;   // the last pass keeps its keys for bA
            uint32_t *vout = last ? d_iA : (p % 2 == 0 ? v1 : v2);
            k_radix_hist<<<(unsigned)ntiles, kBT, 0, st>>>(kin, nions, p * kRB, dtc, (uint32_t)ntiles);
            HCX_HIP(hipGetLastError());
            rc = scan_exclusive(st, dtc, dtc, tc_words, dtmp);
            if (rc) return rc;
            k_radix_scatter<<<(unsigned)ntiles, kBT, 0, st>>>(kin, vin, kout, vout, nions, p * kRB, dtc,
                                                                (uint32_t)ntiles);
            HCX_HIP(hipGetLastError());
            kin = kout;
            vin = vout;
        }
        sorted = kin;
    }
    k_bucket_offsets<<<(NB + 1 + 255) / 256, 256, 0, st>>>(sorted, nions, NB, d_bA);
    HCX_HIP(hipGetLastError());
    const uint32_t D = max_mass + 2;
    k_mdir<<<(D + 255) / 256, 256, 0, st>>>(d_mass, a.n, d_mdir, D);
    HCX_HIP(hipGetLastError());
    // the scratch stays with the context: later builds on this stream reuse it in order
    return HCX_SUCCESS;
}

}  // namespace hcx
