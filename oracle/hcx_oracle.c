/*
 * hcx_oracle.c -- CPU restatement of HiCOPS' one-rank search core.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the reported
 * CPU baseline.  Nothing in the product library links or calls this file.
 *
 * Each function follows the reference at the file:line in its comment
 * (paths relative to the reference repository).  Floating-point expressions
 * keep the reference's types and evaluation order; build with
 * -ffp-contract=off so no multiply-add is fused.
 *
 * Pinning (how each function is checked, since the reference itself needs
 * CMake-generated headers and is not built here):
 *   orc_lgfact            closed form lgamma(n+1)/ln 10
 *   orc_pep_mass          published monoisotopic peptide masses (+ residue table values)
 *   orc_mod_mass          unmodified mass + first mod mass (Q4), float arithmetic
 *   orc_gen_*spectrum     hand-traced exact bins of utils.cpp:231-276 / :378-505: full
 *                         b/y ladders at z = 1..3, the b1 of every residue, one- and
 *                         two-mod-type ladders; planted mistakes (ORC_MUTANT_*) are
 *                         rejected by them; textbook b/y m/z within the truncation
 *                         bound; b/y complementarity
 *   orc_index_build       bA prefix-sum invariants, iA permutation + per-bin order,
 *                         brute-force stable sort on tiny inputs
 *   orc_window            brute-force linear scan
 *   orc_search (counts)   brute force: direct comparison of query bins with each
 *                         candidate's own ion bins on tiny inputs
 *   orc_search (score)    closed-form hyperscore recomputed from the counts
 *   orc_heap_replay       hand-traced insert sequences of minheap.h (ties included)
 *   orc_tailfit           textbook OLS on the selected points; hand-traced marks;
 *                         survival function by definition
 */
#include "hcx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* utils.cpp:29-30 */
#define ORC_H2O    18.015f
#define ORC_PROTON 1.00727647
#define ORC_NAA    (-20000)

/* utils.cpp:41-68 (float table, 'A'..'Z'; NAA for non-residues) */
static const float kAAMass[26] = {
    71.03712f, ORC_NAA, 103.00919f, 115.030f, 129.0426f, 147.068f, 57.02146f,
    137.060f, 113.084f, ORC_NAA, 128.094f, 113.084f, 131.0405f, 114.043f, ORC_NAA,
    97.0527f, 128.05858f, 156.1012f, 87.032f, 101.0476f, ORC_NAA, 99.06841f,
    186.0793f, ORC_NAA, 163.0633f, ORC_NAA};
/* utils.cpp:71-98: only C carries a static mod (+57.021464). */
static float stat_mod_of(char c) { return c == 'C' ? 57.021464f : 0.0f; }

float orc_aa_mass(char c) { return kAAMass[c - 'A']; }
float orc_stat_mod(char c) { return stat_mod_of(c); }

/* Planted mistakes, for the pin tests only: tests/test_oracle_pins.py
 * rebuilds this file with one ORC_MUTANT_* macro defined and checks that
 * the hand-traced bins reject the result.  The oracle built by the Makefile
 * defines none of them.
 *   ORC_MUTANT_FLOAT_PROTON    PROTON as a float constant: GETAA and the ion
 *                              arithmetic built on it in float
 *   ORC_MUTANT_SUM_THEN_TRUNC  unmodified ions: one truncation of the running
 *                              sum instead of one per residue term
 *   ORC_MUTANT_MOD_PER_TERM    modified ions: each term divided by the charge
 *                              before the sum (the unmodified rule)
 *   ORC_MUTANT_Y_MOD_ORDER     y-ion mods taken in N-to-C nibble order
 *   ORC_MUTANT_NO_STATMOD      the static C+57.021464 left out */

/* utils.cpp:101 GETAA: float + float, then + double */
#ifdef ORC_MUTANT_FLOAT_PROTON
typedef float getaa_t;   /* a float PROTON makes GETAA and every product using it float */
#else
typedef double getaa_t;
#endif
static getaa_t getaa(char c, uint32_t z)
{
#ifdef ORC_MUTANT_NO_STATMOD
    float f = kAAMass[c - 'A'];
#else
    float f = kAAMass[c - 'A'] + stat_mod_of(c);
#endif
#ifdef ORC_MUTANT_FLOAT_PROTON
    return f + (float)ORC_PROTON * (float)z;
#else
    return (double)f + ORC_PROTON * (double)z;
#endif
}

void orc_lgfact(double *out, uint32_t n)
{
    /* utils.h:115-131 */
    if (n == 0) return;
    out[0] = 0.0;
    if (n > 1) out[1] = 0.0;
    for (uint32_t i = 2; i < n; i++) out[i] = log10((double)i) + out[i - 1];
}

float orc_pep_mass(const char *seq, uint32_t len)
{
    /* utils.cpp:290-303 */
    float mass = ORC_H2O;
    for (uint32_t l = 0; l < len; l++) {
        mass += kAAMass[seq[l] - 'A'];
        mass += stat_mod_of(seq[l]);
    }
    return mass;
}

float orc_mod_mass(const char *seq, uint32_t len, uint32_t mod_num, const hcx_params *p)
{
    /* utils.cpp:337-362.  The loop's second step shifts by the mask value
     * (0xF0), which x86 reduces to a 16-bit shift of an 8-bit field, so only
     * the first mod nibble is ever added (Q4). */
    float mass = orc_pep_mass(seq, len);
    uint32_t first = mod_num & 0x0Fu;
    if (first != 0)
        mass += (float)p->mod_mass[first - 1] / (float)p->scale;
    return mass;
}

float orc_gen_spectrum(const char *seq, uint32_t len, const hcx_params *p, uint32_t *S)
{
    /* utils.cpp:231-276 */
    float mass = orc_pep_mass(seq, len);
    const uint32_t maxz = p->maxz, scale = p->scale;
    if (mass > 0) {
        memset(S, 0, sizeof(uint32_t) * 2 * maxz * (len - 1));
        for (uint32_t z = 0; z < maxz; z++) {
            uint32_t bstart = z * (len - 1);
            uint32_t ystart = z * (len - 1) + maxz * (len - 1);
#ifdef ORC_MUTANT_SUM_THEN_TRUNC
            double bsum = getaa(seq[0], z + 1) * scale, ysum = (getaa(seq[len - 1], z + 1) + ORC_H2O) * scale;
            S[bstart] = (uint32_t)(bsum / (z + 1));
            S[ystart] = (uint32_t)(ysum / (z + 1));
            for (uint32_t l = 1; l < len - 1; l++) {
                bsum += getaa(seq[l], 0) * scale;
                ysum += getaa(seq[len - 1 - l], 0) * scale;
                S[bstart + l] = (uint32_t)(bsum / (z + 1));
                S[ystart + l] = (uint32_t)(ysum / (z + 1));
            }
#else
            S[bstart] = (uint32_t)((getaa(seq[0], z + 1) * scale) / (z + 1));
            S[ystart] = (uint32_t)(((getaa(seq[len - 1], z + 1) + ORC_H2O) * scale) / (z + 1));
            for (uint32_t l = 1; l < len - 1; l++) {
                S[bstart + l] = S[bstart + l - 1] + (uint32_t)((getaa(seq[l], 0) * scale) / (z + 1));
                S[ystart + l] = S[ystart + l - 1] +
                                (uint32_t)((getaa(seq[len - 1 - l], 0) * scale) / (z + 1));
            }
#endif
        }
    }
    return mass;
}

float orc_gen_mod_spectrum(const char *seq, uint32_t len, const hcx_pep_entry *e,
                           const hcx_params *p, uint32_t *S)
{
    /* utils.cpp:378-505 */
    const double minmass = p->min_mass, maxmass = p->max_mass;
    const uint32_t maxz = p->maxz, scale = p->scale;
    float mass;
    char modPos[64];
    int modNums[HCX_MAX_MOD_TYPES];
    int modSeen = 0;

    if (e->sites == 0 || e->mod_num == 0)
        mass = ORC_NAA;
    else
        mass = orc_mod_mass(seq, len, e->mod_num, p);

    uint32_t bitmask = e->mod_num;
    if (mass > minmass && mass < maxmass) {
        for (int i = 0; i < HCX_MAX_MOD_TYPES; i++) {
            modNums[i] = (int)(bitmask & 0x0F) - 1;
            bitmask /= 16;
            if (modNums[i] != -1) modSeen++;
        }
        for (uint32_t i = 0; i < len; i++) modPos[i] = (e->sites >> i) & 1ull ? 1 : 0;

        for (uint32_t z = 0; z < maxz; z++) {
            uint32_t bstart = z * (len - 1);
            uint32_t ystart = z * (len - 1) + maxz * (len - 1);
            S[bstart] = (uint32_t)(getaa(seq[0], z + 1) * scale);
            S[ystart] = (uint32_t)((getaa(seq[len - 1], z + 1) + ORC_H2O) * scale);
            for (uint32_t l = 1; l < len - 1; l++) {
                S[bstart + l] = S[bstart + l - 1] + (uint32_t)(getaa(seq[l], 0) * scale);
                S[ystart + l] = S[ystart + l - 1] + (uint32_t)(getaa(seq[len - 1 - l], 0) * scale);
            }
        }
        for (uint32_t z = 0; z < maxz; z++) {
            uint32_t bstart = z * (len - 1), counter = 0;
            for (uint32_t l = 0; l < len - 1; l++) {
                counter += modPos[l];
#ifdef ORC_MUTANT_MOD_PER_TERM
                uint32_t t = 0;
                for (uint32_t k = 0; k < counter; k++) t += p->mod_mass[modNums[k]] / (z + 1);
                S[bstart + l] = S[bstart + l] / (z + 1) + t;
#else
                for (uint32_t k = 0; k < counter; k++) S[bstart + l] += p->mod_mass[modNums[k]];
                S[bstart + l] /= (z + 1);
#endif
            }
        }
        for (uint32_t z = 0; z < maxz; z++) {
            uint32_t ystart = z * (len - 1) + maxz * (len - 1), counter = 0;
            for (int l = (int)len - 1; l > 0; l--) {
                counter += modPos[l];
#ifdef ORC_MUTANT_Y_MOD_ORDER
                for (uint32_t k = 0; k < counter; k++) S[ystart + (len - 1) - l] += p->mod_mass[modNums[k]];
#else
                for (uint32_t k = 0; k < counter; k++)
                    S[ystart + (len - 1) - l] += p->mod_mass[modNums[modSeen - 1 - k]];
#endif
                S[ystart + (len - 1) - l] /= (z + 1);
            }
        }
    }
    return mass;
}

int orc_index_build(const hcx_params *p, uint32_t peplen, const char *seqs,
                    const hcx_pep_entry *entries, uint32_t n, uint32_t *iA, uint32_t *bA)
{
    /* dslim.cpp:216-347 (ions + histogram), :89-126 (bA scan), :361-388 +
     * :452-482 (key-value sort, then ascending sort inside each bin: the net
     * result is a stable sort of ascending ion ids by bin). */
    if (peplen < 2 || p->maxz == 0) return HCX_ERR_INVLD_PARAM;
    const uint32_t speclen = 2 * p->maxz * (peplen - 1);
    const uint32_t NB = p->max_mass * p->scale;
    const double minmass = p->min_mass, maxmass = p->max_mass;
    uint64_t total = (uint64_t)n * speclen;
    uint32_t *spec = (uint32_t *)malloc(sizeof(uint32_t) * (total ? total : 1));
    uint32_t *cnt = (uint32_t *)calloc((size_t)NB + 1, sizeof(uint32_t));
    uint32_t *tmp = (uint32_t *)calloc(speclen, sizeof(uint32_t));
    if (!spec || !cnt || !tmp) { free(spec); free(cnt); free(tmp); return HCX_ERR_BAD_MEM_ALLOC; }

    for (uint32_t k = 0; k < n; k++) {
        const hcx_pep_entry *e = entries + k;
        const char *seq = seqs + (size_t)e->seq_id * peplen;
        float m;
        memset(tmp, 0, sizeof(uint32_t) * speclen);
        if (e->mod_num == 0) m = orc_gen_spectrum(seq, peplen, p, tmp);
        else m = orc_gen_mod_spectrum(seq, peplen, e, p, tmp);
        uint32_t *dst = spec + (uint64_t)k * speclen;
        if (m >= minmass && m <= maxmass) {
            for (uint32_t ion = 0; ion < speclen; ion++) {
                uint32_t b = tmp[ion];
                if (b >= maxmass * p->scale) b = (uint32_t)(maxmass * p->scale) - 1;
                dst[ion] = b;
                cnt[b]++;
            }
        } else {
            memset(dst, 0, sizeof(uint32_t) * speclen);
            cnt[0] += speclen;
        }
    }
    bA[0] = 0;
    for (uint32_t i = 1; i <= NB; i++) bA[i] = bA[i - 1] + cnt[i - 1];
    /* stable counting scatter: ascending id inside each bin */
    memcpy(cnt, bA, sizeof(uint32_t) * NB);
    for (uint64_t id = 0; id < total; id++) iA[cnt[spec[id]]++] = (uint32_t)id;
    free(spec); free(cnt); free(tmp);
    return (bA[NB] == total) ? HCX_SUCCESS : HCX_ERR_INVLD_SIZE;
}

/* DSLIM_Construct over several lengths at once (dslim.cpp:46-138 runs one
 * length after another; each is independent): OpenMP over the lengths, for
 * the CPU baseline.  threads <= 0: the OpenMP default. */
int orc_index_build_all(const hcx_params *p, uint32_t nlen, const uint32_t *peplen, const char *const *seqs,
                        const hcx_pep_entry *const *entries, const uint32_t *n, uint32_t *const *iA,
                        uint32_t *const *bA, int threads)
{
    int rc = HCX_SUCCESS;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < (int64_t)nlen; i++) {
        const int r = orc_index_build(p, peplen[i], seqs[i], entries[i], n[i], iA[i], bA[i]);
        if (r != HCX_SUCCESS) {
#ifdef _OPENMP
#pragma omp critical
#endif
            rc = r;
        }
    }
    return rc;
}

/* dslim_query.cpp:919-955 */
static int32_t bin_find_min(const float *m, float pmass1, int32_t min, int32_t max)
{
    for (;;) {
        int32_t half = (min + max) / 2;
        if (max - min < 20) {
            int32_t cur = min;
            while (m[cur] < pmass1) cur++;
            return cur;
        }
        if (pmass1 > m[half]) { min = half; continue; }
        if (pmass1 < m[half]) { max = half; continue; }
        /* equal: walk back to the first equal mass (guarded at 0, where the
         * reference would read entries[-1]) */
        while (half >= 0 && pmass1 == m[half]) half--;
        return half + 1;
    }
}

/* dslim_query.cpp:957-994 */
static int32_t bin_find_max(const float *m, float pmass2, int32_t min, int32_t max, int32_t n)
{
    for (;;) {
        int32_t half = (min + max) / 2;
        if (max - min < 20) {
            int32_t cur = max;
            while (m[cur] > pmass2) cur--;
            return cur;
        }
        if (pmass2 > m[half]) { min = half; continue; }
        if (pmass2 < m[half]) { max = half; continue; }
        half++;
        while (half < n && pmass2 == m[half]) half++;
        return half - 1;
    }
}

int orc_window(const float *mass, uint32_t n, float precursor, double dM,
               int32_t *minl, int32_t *maxl)
{
    /* dslim_query.cpp:854-916 */
    float pmass1 = (float)((double)precursor - dM);
    float pmass2 = (float)((double)precursor + dM);
    if (n == 0) { *minl = 0; *maxl = -1; return 0; }
    int32_t min = 0, max = (int32_t)n - 1;
    if (dM < 0.0) { *minl = min; *maxl = max; return 0; }
    if (pmass1 < mass[min]) *minl = min;
    else if (pmass1 > mass[max]) { *minl = max; *maxl = max; return 0; }
    else *minl = bin_find_min(mass, pmass1, min, max);
    if (pmass2 > mass[max]) *maxl = max;
    else if (pmass2 < mass[min]) { *minl = min; *maxl = min; return 0; }
    else *maxl = bin_find_max(mass, pmass2, min, max, (int32_t)n);
    return (mass[*maxl] <= pmass2 && mass[*minl] >= pmass1) ? 1 : 0;
}

/* ---------------- minheap.h:86-277 ---------------- */
typedef struct orc_cell {
    float hs;
    int32_t psid;
    uint16_t idxoffset, shpk, speclen, pad;
    uint32_t bc, yc;
    int32_t order;   /* insertion order (for orc_heap_replay) */
} orc_cell;

typedef struct orc_heap { int size, cap; orc_cell *a; } orc_heap;

static void heap_heapify(orc_heap *h, int pos)
{
    for (;;) {
        int r = (pos + 1) * 2, l = r - 1, s = pos;
        float sv = h->a[s].hs;
        if (l < h->size && h->a[l].hs < sv) { s = l; sv = h->a[l].hs; }
        if (r < h->size && h->a[r].hs < sv) { s = r; sv = h->a[r].hs; }
        if (s == pos) return;
        orc_cell t = h->a[s]; h->a[s] = h->a[pos]; h->a[pos] = t;
        pos = s;
    }
}

static void heap_insert(orc_heap *h, const orc_cell *e)
{
    if (h->size == h->cap) {                     /* increase_key(0, e) */
        if (e->hs < h->a[0].hs) return;
        h->a[0] = *e;
        heap_heapify(h, 0);
        return;
    }
    h->a[h->size++] = *e;
    for (int i = (h->size - 1) / 2; i >= 0; i--) heap_heapify(h, i);
}

static orc_cell heap_getmax(const orc_heap *h)
{
    if (h->size <= 1) return h->a[0];
    orc_cell m = h->a[h->size - 1];
    for (int k = h->size - 2; k >= h->size / 2; k--)
        if (h->a[k].hs > m.hs) m = h->a[k];
    return m;
}

int orc_heap_capacity(uint32_t topmatches)
{
    /* dslim.cpp:422-428 */
    uint32_t num = topmatches == 0 ? 1 : topmatches;
    num = (uint32_t)floor(log2((double)num) + 0.999);
    return (1 << num) - 1;
}

int orc_heap_replay(const float *hs, uint32_t n, int32_t cap, int32_t *argmax, int32_t *final_size)
{
    if (cap <= 0) return HCX_ERR_INVLD_PARAM;
    orc_heap h = {0, cap, (orc_cell *)calloc((size_t)cap, sizeof(orc_cell))};
    for (uint32_t i = 0; i < n; i++) {
        orc_cell c; memset(&c, 0, sizeof c);
        c.hs = hs[i]; c.order = (int32_t)i;
        heap_insert(&h, &c);
    }
    orc_cell m = heap_getmax(&h);
    *argmax = h.size ? m.order : -1;
    *final_size = h.size;
    free(h.a);
    return HCX_SUCCESS;
}

/* ---------------- expeRT.cpp:572-718 ---------------- */
static int rargmax_d(const double *d, int i1, int i2, double v)
{   /* expeRT.cpp:1208-1223 */
    int rv = i2;
    for (int p = i2; p >= i1; p--) if (d[p] >= v) { rv = p; break; }
    return rv;
}
static int argmax_d(const double *d, int i1, int i2, double v)
{   /* expeRT.cpp:1225-1240 */
    int rv = i1;
    for (int p = i1; p <= i2; p++) if (d[p] >= v) { rv = p; break; }
    return rv;
}
static int largmax_d(const double *d, int i1, int i2, double v)
{   /* expeRT.cpp:1242-1257 */
    int rv = i1;
    for (int p = i1; p <= i2; p++) if (d[p] <= v) { rv = p; break; }
    return rv;
}

int orc_tailfit(const double *yy, uint32_t hbins, uint32_t cpsms, int32_t hyp,
                int32_t *mu, int32_t *beta, int32_t *minhyp, int32_t *nexthyp, uint32_t *flags)
{
    double mu_t = 0.0, beta_t = 4.0;
    uint32_t fl = 0;
    if (hyp < 1 || hyp >= (int32_t)hbins) {
        /* the reference reads yy[-1] or past the histogram (UB): defined here
         * as a zero fit */
        *mu = 0; *beta = 0; *minhyp = 0; *nexthyp = 0;
        if (flags) *flags |= HCX_RES_FIT_UB;
        return HCX_SUCCESS;
    }
    int end1 = rargmax_d(yy, 0, hyp - 1, 1.0);
    int stt1 = argmax_d(yy, 0, end1, 1.0);
    if (stt1 == end1) end1 += 1;
    int n = end1 - stt1 + 1;
    int vaa = (int)cpsms;
    if (hyp < 1 || n < 2 || (uint32_t)end1 >= hbins) {
        /* reference reads outside the histogram or the slice (UB) */
        fl |= HCX_RES_FIT_UB;
        mu_t = 0.0; beta_t = 0.0;
    } else if (vaa < 1) {
        mu_t = 0.0; beta_t = 100.0;
    } else {
        double *sx = (double *)malloc(sizeof(double) * (size_t)n);
        double acc = 0.0;
        for (int i = 0; i < n; i++) {          /* partial_sum, divide, add(-1), negative */
            acc = acc + yy[stt1 + i];
            double t = acc / (double)vaa;
            t = t + (-1.0);
            sx[i] = -t;
        }
        for (int i = 0; i < n; i++) if (sx[i] > 1.0) sx[i] = 0.999;
        int rep = rargmax_d(sx, 0, n - 1, 1e-4);
        double rv = sx[rep];
        for (int i = 0; i < n; i++) if (sx[i] <= 0.0) sx[i] = rv;
        for (int i = 0; i < n; i++) sx[i] = log10(sx[i]);
        int mark = 0, mark2 = 0;
        double hgt = sx[n - 1] - sx[0];
        if (n > 3) {
            mark = largmax_d(sx, 0, n - 1, sx[0] + hgt * 0.22) - 1;
            mark2 = rargmax_d(sx, 0, n - 1, sx[0] + hgt * 0.87);
            if (mark2 == n) mark2 -= 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else if (n == 3) {
            mark = largmax_d(sx, 0, n - 1, sx[0] + hgt * 0.22) - 1;
            mark2 = n - 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else {
            mark = 0;
            mark2 = n - 1;
        }
        if (mark < 0) {
            /* the reference clips to head-1 and reads sx[-1] (Q20) */
            fl |= HCX_RES_FIT_UB;
            mu_t = 0.0; beta_t = 0.0;
        } else {
            int m = mark2 - mark + 1;
            const double *y = sx + mark;
            if (m == 1) {                         /* expeRT.cpp:1308-1313 */
                mu_t = 0.0; beta_t = y[0];
            } else {                              /* expeRT.cpp:1317-1342 */
                double xbar = 0.0, ybar = 0.0, top = 0.0, bot = 0.0;
                for (int i = 0; i < m; i++) {
                    xbar = xbar + (double)(stt1 + mark + i);
                    ybar = ybar + y[i];
                }
                xbar = xbar / (double)m;
                ybar = ybar / (double)m;
                for (int i = 0; i < m; i++) {
                    double dx = (double)(stt1 + mark + i) - xbar;
                    top = top + dx * (y[i] - ybar);
                    bot = bot + dx * dx;
                }
                mu_t = top / bot;
                beta_t = ybar - mu_t * xbar;
            }
            if (!(fabs(mu_t) < 2.0e3) || !(fabs(beta_t) < 2.0e3)) {
                fl |= HCX_RES_FIT_UB;   /* out of int range after *1e6 */
                mu_t = 0.0; beta_t = 0.0;
            }
        }
        free(sx);
    }
    *mu = (int32_t)(mu_t * 1e6);
    *beta = (int32_t)(beta_t * 1e6);
    *minhyp = stt1;
    *nexthyp = end1;
    if (flags) *flags |= fl;
    return HCX_SUCCESS;
}

/* ---------------- multi-rank partial results and merge ---------------- */
/* The reference's (ushort) conversions of doubles outside 0..65535 and its
 * int product k * 65500 are undefined in C++; x86-64 g++ code converts
 * through a 32-bit cvttsd2si and keeps the low 16 bits, and multiplies in
 * 32 bits with wraparound.  Both are restated that way. */
static uint16_t u16_of(double v) { return (uint16_t)(uint32_t)(int32_t)v; }

/* dslim_query.cpp:733-763 and expeRT::StoreIResults (expeRT.cpp:870-906) */
int orc_part_export(const hcx_result *res, const uint32_t *hist, uint32_t hbins, uint32_t nspec, int32_t qid0,
                    uint32_t slot, hcx_part_res *packs, uint16_t *slices)
{
    for (uint32_t s = 0; s < nspec; s++) {
        const hcx_result *r = res + s;
        hcx_part_res *pr = packs + s;
        uint16_t *out = slices + (size_t)s * slot;
        memset(pr, 0, sizeof *pr);
        memset(out, 0, sizeof(uint16_t) * slot);
        if ((int32_t)r->cpsms >= 1) {
            const uint32_t *yy = hist + (size_t)s * hbins;
            int ends = (int)hbins - 1, stt = 0;          /* rargmax / argmax at 0.99 */
            for (int p = (int)hbins - 1; p >= 0; p--) if ((double)yy[p] >= 0.99) { ends = p; break; }
            for (int p = 0; p <= ends; p++) if ((double)yy[p] >= 0.99) { stt = p; break; }
            if ((uint32_t)(ends - stt + 1) > slot) return HCX_ERR_INVLD_SIZE;
            for (int ii = stt; ii <= ends; ii++) {
                uint16_t k = u16_of((double)yy[ii]);
                if ((int32_t)r->cpsms > 65500)
                    k = u16_of((double)(int32_t)((uint32_t)k * 65500u) / (double)(int32_t)r->cpsms);
                out[ii - stt] = k;
            }
            pr->min = (uint16_t)stt;
            pr->max2 = (uint16_t)ends;
            pr->max = r->hyperscore;
            pr->N = (int32_t)r->cpsms;
        }
        pr->qID = qid0 + (int32_t)s;
    }
    return HCX_SUCCESS;
}

/* expeRT::ModelTailFit(double&, int) (expeRT.cpp:720-852) on the merged
 * histogram pdata with pN candidates: unlike the one-rank fit, no widening
 * of a one-bin range, vaa2 = max(std::accumulate(range, 1) in an int, pN),
 * the counts divided before the running sum, and the slope and intercept
 * used unrounded.  Undefined paths of the reference (hyp outside the
 * histogram, mark < 0) are a zero line, flagged. */
double orc_tailfit_merged(const double *pdata, uint32_t hbins, int32_t pN, int32_t hyp, uint32_t *flags)
{
    double mu_t = 0.0, beta_t = 4.0;
    if (hyp < 1 || hyp >= (int32_t)hbins) {
        /* the reference reads outside pdata: not reported */
        if (flags) *flags |= HCX_RES_FIT_UB;
        return INFINITY;
    }
    int end1 = rargmax_d(pdata, 0, hyp - 1, 1.0);
    int stt1 = argmax_d(pdata, 0, end1, 1.0);
    int n = end1 - stt1 + 1;
    int vaa = pN;
    if (vaa < 1) {
        mu_t = 0.0; beta_t = 100.0;
    } else {
        int acc = 1;
        for (int i = 0; i < n; i++) acc = (int)((double)acc + pdata[stt1 + i]);
        const int vaa2 = acc > vaa ? acc : vaa;
        double *sx = (double *)malloc(sizeof(double) * (size_t)n);
        double run = 0.0;
        for (int i = 0; i < n; i++) {            /* divide, partial_sum, add(-1), negative */
            run = (i == 0) ? pdata[stt1] / (double)vaa2 : run + pdata[stt1 + i] / (double)vaa2;
            double t = run + (-1.0);
            sx[i] = -t;
        }
        for (int i = 0; i < n; i++) if (sx[i] > 1.0) sx[i] = 0.999;
        int rep = rargmax_d(sx, 0, n - 1, 1e-4);
        double rv = sx[rep];
        for (int i = 0; i < n; i++) if (sx[i] <= 0.0) sx[i] = rv;
        for (int i = 0; i < n; i++) sx[i] = log10(sx[i]);
        int mark = 0, mark2 = 0;
        double hgt = sx[n - 1] - sx[0];
        if (n > 3) {
            mark = largmax_d(sx, 0, n - 1, sx[0] + hgt * 0.22) - 1;
            mark2 = rargmax_d(sx, 0, n - 1, sx[0] + hgt * 0.87);
            if (mark2 == n) mark2 -= 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else if (n == 3) {
            mark = largmax_d(sx, 0, n - 1, sx[0] + hgt * 0.22) - 1;
            mark2 = n - 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else {
            mark = 0;
            mark2 = n - 1;
        }
        if (mark < 0) {
            if (flags) *flags |= HCX_RES_FIT_UB;
            mu_t = 0.0; beta_t = 0.0;
        } else {
            int m = mark2 - mark + 1;
            const double *y = sx + mark;
            if (m == 1) {                         /* expeRT.cpp:1308-1313 */
                mu_t = 0.0; beta_t = y[0];
            } else {                              /* expeRT.cpp:1317-1342 */
                double xbar = 0.0, ybar = 0.0, top = 0.0, bot = 0.0;
                for (int i = 0; i < m; i++) {
                    xbar = xbar + (double)(stt1 + mark + i);
                    ybar = ybar + y[i];
                }
                xbar = xbar / (double)m;
                ybar = ybar / (double)m;
                for (int i = 0; i < m; i++) {
                    double dx = (double)(stt1 + mark + i) - xbar;
                    top = top + dx * (y[i] - ybar);
                    bot = bot + dx * dx;
                }
                mu_t = top / bot;
                beta_t = ybar - mu_t * xbar;
            }
        }
        free(sx);
    }
    return pow(10, hyp * mu_t + beta_t) * vaa;
}

/* DSLIM_Score::CombineResults, dslim_score.cpp:244-345, with
 * expeRT::Reconstruct (expeRT.cpp:909-936).  eValue = e * 1e6 into an int
 * (dslim_score.cpp:300, 318): out of the int range that is undefined in C++;
 * x86-64's cvttsd2si gives INT_MIN, restated explicitly. */
static int32_t i32_of(double v) { return (v > -2147483649.0 && v < 2147483648.0) ? (int32_t)v : INT32_MIN; }
int orc_combine(const hcx_params *p, uint32_t nshards, uint32_t nspec, uint32_t slot,
                const hcx_part_res *const *packs, const uint16_t *const *slices, hcx_fresult *out, int32_t *keys)
{
    const uint32_t hbins = 10 * p->max_hyperscore + 2;
    double *pdata = (double *)malloc(sizeof(double) * hbins);
    if (!pdata) return HCX_ERR_BAD_MEM_ALLOC;
    for (uint32_t spec = 0; spec < nspec; spec++) {
        memset(pdata, 0, sizeof(double) * hbins);
        int32_t cpsms = 0, pN = 0, key = (int32_t)nshards;
        float maxhyp = -1;
        for (uint32_t sno = 0; sno < nshards; sno++) {
            const hcx_part_res *r = packs[sno] + spec;
            if (r->min == 0 && r->max == 0 && r->N == 0 && r->max2 == 0) continue;   /* *sResult == 0 */
            cpsms += r->N;
            if (r->N >= 1) {
                const uint16_t *buf = slices[sno] + (size_t)spec * slot;
                pN += r->N;
                for (int jj = r->min; jj <= (int)r->max2; jj++) {
                    double v = buf[jj - r->min];
                    if (r->N > 65500) v = (v / 65500) * r->N;
                    pdata[jj] = pdata[jj] + v;
                }
                if (r->max > 0 && r->max > maxhyp) { maxhyp = r->max; key = (int32_t)sno; }
            }
        }
        hcx_fresult *f = out + spec;
        int reported = 0;
        if (key < (int32_t)nshards && cpsms >= (int32_t)p->min_cpsm) {
            const int32_t hyp = (int32_t)((double)(maxhyp * 10.0f) + 0.5);
            const double e_x = orc_tailfit_merged(pdata, hbins, pN, hyp, NULL);
            if (e_x < p->expect_max) {
                f->eValue = i32_of(e_x * 1e6);
                f->specID = packs[key][spec].qID;
                f->npsms = cpsms;
                keys[spec] = key;
                reported = 1;
            }
        }
        if (!reported) {
            f->eValue = i32_of(p->expect_max * 1e6);
            f->specID = -1;
            f->npsms = 0;
            keys[spec] = (int32_t)nshards;
        }
    }
    free(pdata);
    return HCX_SUCCESS;
}

/* ---------------- dslim_query.cpp:515-842 ---------------- */
typedef struct orc_byc { int16_t bc, yc; int64_t ibc, iyc; } orc_byc;

typedef struct orc_ws {
    orc_byc *byc;
    double *surv;
    orc_heap heap;
} orc_ws;

/* One spectrum.  cand (optional) receives every inserted candidate. */
static void search_one(const hcx_params *p, const orc_index *idx, int nidx,
                       const hcx_queries_view *q, int32_t s, const double *lgf, uint32_t hbins,
                       orc_ws *ws, hcx_result *res, uint32_t *hist, hcx_stats *st,
                       hcx_cand *cand, uint64_t cap, uint64_t *ncand)
{
    const uint32_t *moz = q->moz + q->idx[s];
    const uint32_t *inten = q->intensity + q->idx[s];
    const uint32_t np = q->idx[s + 1] - q->idx[s];
    const float pmass = q->precurse[s];
    const uint32_t dF = p->dF;
    const double maxmass = p->max_mass;
    const uint32_t scale = p->scale;
    uint32_t cpsms = 0;
    uint16_t flags = 0;
    uint64_t n_ia = 0, n_probe = 0;

    memset(ws->surv, 0, sizeof(double) * hbins);
    ws->heap.size = 0;
    memset(ws->heap.a, 0, sizeof(orc_cell) * (size_t)ws->heap.cap);

    for (int ixx = 0; ixx < nidx; ixx++) {
        const orc_index *I = idx + ixx;
        const uint32_t speclen = (I->peplen - 1) * p->maxz * 2;
        const uint32_t half = speclen / 2;
        int32_t minl = 0, maxl = 0;
        int val = orc_window(I->mass, I->n, pmass, p->dM, &minl, &maxl);
        if (!val || maxl < minl) continue;
        orc_byc *byc = ws->byc;
        for (uint32_t k = 0; k < np; k++) {
            uint32_t qion = moz[k], intn = inten[k];
            if (qion > dF && qion < (maxmass * scale) - 1 - dF) {
                for (uint32_t bin = qion - dF; bin < qion + 1 + dF; bin++) {
                    uint32_t start = I->bA[bin], end = I->bA[bin + 1];
                    n_probe++;
                    if (end - start < 1) continue;
                    uint32_t lo_key = (uint32_t)minl * speclen;
                    uint32_t hi_key = ((uint32_t)(maxl + 1) * speclen) - 1;
                    /* std::lower_bound(start, end, lo_key) */
                    uint32_t a = start, b = end;
                    while (a < b) { uint32_t m = a + (b - a) / 2; if (I->iA[m] < lo_key) a = m + 1; else b = m; }
                    uint32_t stt = a;
                    /* std::upper_bound(stt, end, hi_key) */
                    b = end;
                    while (a < b) { uint32_t m = a + (b - a) / 2; if (hi_key < I->iA[m]) b = m; else a = m + 1; }
                    uint32_t ends = a;   /* one past the last */
                    for (uint32_t ion = stt; ion < ends; ion++) {
                        uint32_t raw = I->iA[ion];
                        int32_t ppid = (int32_t)(raw / speclen);
                        int32_t residue = (int32_t)(raw % speclen);
                        int32_t isY = residue / (int32_t)half;
                        int32_t isB = 1 - isY;
                        orc_byc *el = byc + ppid;
                        el->bc += isB;
                        el->ibc += intn * (uint32_t)isB;
                        el->yc += isY;
                        el->iyc += intn * (uint32_t)isY;
                    }
                    n_ia += ends - stt;
                }
            }
        }
        /* dslim_query.cpp:679-725 */
        for (int32_t it = minl; it <= maxl; it++) {
            uint16_t bcc = (uint16_t)byc[it].bc;
            uint16_t ycc = (uint16_t)byc[it].yc;
            uint16_t shpk = (uint16_t)(bcc + ycc);
            if (shpk >= p->min_shp) {
                uint32_t bi = bcc, yi = ycc;
                if (bi > p->max_shdpeaks) { bi = p->max_shdpeaks; flags |= HCX_RES_LGF_CLAMP; }
                if (yi > p->max_shdpeaks) { yi = p->max_shdpeaks; flags |= HCX_RES_LGF_CLAMP; }
                double h1 = lgf[bi] + lgf[yi];
                float hs = (float)(h1 + log10((double)(1 + byc[it].ibc)) +
                                   log10((double)(1 + byc[it].iyc)) - 6);
                if (hs > 0) {
                    orc_cell c;
                    memset(&c, 0, sizeof c);
                    c.hs = hs; c.psid = it; c.idxoffset = (uint16_t)ixx; c.shpk = shpk;
                    c.speclen = (uint16_t)speclen; c.bc = bcc; c.yc = ycc;
                    heap_insert(&ws->heap, &c);
                    cpsms += 1;
                    int bin = (int)(hs * 10 + 0.5);
                    if (bin >= (int)hbins) { bin = (int)hbins - 1; flags |= HCX_RES_HIST_CLAMP; }
                    ws->surv[bin] += 1;
                    if (cand && *ncand < cap) {
                        hcx_cand *o = cand + *ncand;
                        memset(o, 0, sizeof *o);
                        o->spec = (uint32_t)s; o->idxoffset = (uint16_t)ixx; o->psid = it;
                        o->bc = bcc; o->yc = ycc; o->ibc = (uint64_t)byc[it].ibc;
                        o->iyc = (uint64_t)byc[it].iyc; o->hyperscore = hs;
                    }
                    if (cand) (*ncand)++;
                }
            }
        }
        memset(byc + minl, 0, sizeof(orc_byc) * (size_t)(maxl - minl + 1));
    }

    hcx_result r;
    memset(&r, 0, sizeof r);
    r.cpsms = cpsms;
    if (cpsms >= p->min_cpsm && ws->heap.size > 0) {
        orc_cell psm = heap_getmax(&ws->heap);
        r.hyperscore = psm.hs; r.psid = psm.psid; r.idxoffset = psm.idxoffset;
        r.sharedions = psm.shpk; r.totalions = psm.speclen; r.bc = psm.bc; r.yc = psm.yc;
        r.maxhypscore = (int32_t)(psm.hs * 10 + 0.5);
        uint32_t fl = 0;
        orc_tailfit(ws->surv, hbins, cpsms, r.maxhypscore, &r.mu, &r.beta, &r.minhypscore,
                    &r.nexthypscore, &fl);
        flags |= (uint16_t)fl;
        double w = r.mu, b = r.beta;
        w /= 1e6;
        b /= 1e6;
        double lgs_x = (w * r.maxhypscore) + b;
        double e_x = pow(10, lgs_x);
        e_x *= cpsms;
        r.e_x = e_x;
        flags |= HCX_RES_VALID;
        if (e_x < p->expect_max) flags |= HCX_RES_REPORTED;
    }
    r.flags = flags;
    if (res) res[s] = r;
    if (hist) for (uint32_t i = 0; i < hbins; i++) hist[(uint64_t)s * hbins + i] = (uint32_t)ws->surv[i];
    if (st) { st[s].n_ia = n_ia; st[s].n_probe = n_probe; }
}

static uint32_t max_entries(const orc_index *idx, int nidx)
{
    uint32_t m = 1;
    for (int i = 0; i < nidx; i++) if (idx[i].n > m) m = idx[i].n;
    return m;
}

static int ws_init(orc_ws *ws, uint32_t nmax, uint32_t hbins, int cap)
{
    ws->byc = (orc_byc *)calloc(nmax, sizeof(orc_byc));
    ws->surv = (double *)calloc(hbins, sizeof(double));
    ws->heap.cap = cap; ws->heap.size = 0;
    ws->heap.a = (orc_cell *)calloc((size_t)cap, sizeof(orc_cell));
    return (ws->byc && ws->surv && ws->heap.a) ? 0 : -1;
}
static void ws_free(orc_ws *ws) { free(ws->byc); free(ws->surv); free(ws->heap.a); }

int orc_search(const hcx_params *p, const orc_index *idx, int nidx,
               const hcx_queries_view *q, hcx_result *res, uint32_t *hist,
               hcx_stats *stats, int threads)
{
    int cap = orc_heap_capacity(p->topmatches);
    if (cap <= 0) return HCX_ERR_INVLD_PARAM;
    const uint32_t hbins = 10 * p->max_hyperscore + 2;
    double *lgf = (double *)malloc(sizeof(double) * (p->max_shdpeaks + 1));
    orc_lgfact(lgf, p->max_shdpeaks + 1);
    const uint32_t nmax = max_entries(idx, nidx);
    int status = HCX_SUCCESS;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        orc_ws ws;
        if (ws_init(&ws, nmax, hbins, cap) != 0) {
            status = HCX_ERR_BAD_MEM_ALLOC;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
            for (int32_t s = 0; s < q->num_specs; s++)
                search_one(p, idx, nidx, q, s, lgf, hbins, &ws, res, hist, stats, NULL, 0, NULL);
        }
        ws_free(&ws);
    }
    free(lgf);
    return status;
}

int orc_candidates(const hcx_params *p, const orc_index *idx, int nidx,
                   const hcx_queries_view *q, int32_t s, hcx_cand *out, uint64_t cap, uint64_t *n)
{
    int hcap = orc_heap_capacity(p->topmatches);
    if (hcap <= 0) return HCX_ERR_INVLD_PARAM;
    const uint32_t hbins = 10 * p->max_hyperscore + 2;
    double *lgf = (double *)malloc(sizeof(double) * (p->max_shdpeaks + 1));
    orc_lgfact(lgf, p->max_shdpeaks + 1);
    orc_ws ws;
    if (ws_init(&ws, max_entries(idx, nidx), hbins, hcap) != 0) { ws_free(&ws); free(lgf); return HCX_ERR_BAD_MEM_ALLOC; }
    *n = 0;
    search_one(p, idx, nidx, q, s, lgf, hbins, &ws, NULL, NULL, NULL, out, cap, n);
    ws_free(&ws);
    free(lgf);
    return HCX_SUCCESS;
}

int orc_count_work(const hcx_params *p, const orc_index *idx, int nidx,
                   const hcx_queries_view *q, hcx_stats *stats, int threads)
{
    const double maxmass = p->max_mass;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16)
#endif
    for (int32_t s = 0; s < q->num_specs; s++) {
        const uint32_t *moz = q->moz + q->idx[s];
        const uint32_t np = q->idx[s + 1] - q->idx[s];
        uint64_t n_ia = 0, n_probe = 0;
        for (int ixx = 0; ixx < nidx; ixx++) {
            const orc_index *I = idx + ixx;
            const uint32_t speclen = (I->peplen - 1) * p->maxz * 2;
            int32_t minl = 0, maxl = 0;
            if (!orc_window(I->mass, I->n, q->precurse[s], p->dM, &minl, &maxl) || maxl < minl) continue;
            uint32_t lo_key = (uint32_t)minl * speclen, hi_key = ((uint32_t)(maxl + 1) * speclen) - 1;
            for (uint32_t k = 0; k < np; k++) {
                uint32_t qion = moz[k];
                if (!(qion > p->dF && qion < (maxmass * p->scale) - 1 - p->dF)) continue;
                for (uint32_t bin = qion - p->dF; bin < qion + 1 + p->dF; bin++) {
                    uint32_t start = I->bA[bin], end = I->bA[bin + 1];
                    n_probe++;
                    if (end == start) continue;
                    uint32_t a = start, b = end;
                    while (a < b) { uint32_t m = a + (b - a) / 2; if (I->iA[m] < lo_key) a = m + 1; else b = m; }
                    uint32_t stt = a;
                    b = end;
                    while (a < b) { uint32_t m = a + (b - a) / 2; if (hi_key < I->iA[m]) b = m; else a = m + 1; }
                    n_ia += a - stt;
                }
            }
        }
        stats[s].n_ia = n_ia;
        stats[s].n_probe = n_probe;
    }
    return HCX_SUCCESS;
}

/* ---- Queries producer: keyval.h:48-132, msquery.cpp:509-569 ---------------- */

static void keyval_rec(uint32_t *key, uint32_t *val, int64_t left, int64_t right)
{
    /* keyval.h:50-80 partition */
    int64_t i = left, j = right;
    const uint32_t pivot = key[(left + right) / 2];
    while (i <= j) {
        while (key[i] < pivot) i++;
        while (key[j] > pivot) j--;
        if (i <= j) {
            uint32_t tk = key[i], tv = val[i];
            key[i] = key[j];
            val[i] = val[j];
            key[j] = tk;
            val[j] = tv;
            i++;
            j--;
        }
    }
    /* keyval.h:82-105: both halves, serially or as tasks */
    if (left < j) keyval_rec(key, val, left, j);
    if (i < right) keyval_rec(key, val, i, right);
}

void orc_keyval_sort(uint32_t *key, uint32_t *val, uint32_t n)
{
    /* keyval.h:124-127 sorts [0, lenArray-1]; with lenArray 0 the partition
     * loop never runs */
    if (n > 0) keyval_rec(key, val, 0, (int64_t)n - 1);
}

int orc_pickpeaks(const uint32_t *raw_mz, const uint32_t *raw_in, const uint32_t *raw_idx,
                  const double *prec_mz, const int32_t *z, const double *rtime, int32_t n,
                  uint32_t qalen, uint32_t base_int, uint32_t min_int,
                  uint32_t *moz, uint32_t *intensity, uint32_t *idx,
                  float *precurse, int32_t *charges, float *rtimes)
{
    if (n < 0 || qalen == 0) return HCX_ERR_INVLD_PARAM;
    uint32_t maxS = 0;
    for (int32_t s = 0; s < n; s++) {
        if (raw_idx[s + 1] < raw_idx[s]) return HCX_ERR_INVLD_PARAM;
        if (raw_idx[s + 1] - raw_idx[s] > maxS) maxS = raw_idx[s + 1] - raw_idx[s];
    }
    uint32_t *key = (uint32_t *)malloc((size_t)(maxS ? maxS : 1) * 4);
    uint32_t *val = (uint32_t *)malloc((size_t)(maxS ? maxS : 1) * 4);
    if (!key || !val) { free(key); free(val); return HCX_ERR_BAD_MEM_ALLOC; }
    idx[0] = 0;
    for (int32_t s = 0; s < n; s++) {
        const int64_t S = raw_idx[s + 1] - raw_idx[s];
        memcpy(key, raw_in + raw_idx[s], (size_t)S * 4);
        memcpy(val, raw_mz + raw_idx[s], (size_t)S * 4);
        /* msquery.cpp:515-517 */
        precurse[s] = (float)prec_mz[s];
        charges[s] = 1 > z[s] ? 1 : z[s];
        {
            double rt = rtime ? rtime[s] : 0.0;
            rtimes[s] = (float)(0.0 > rt ? 0.0 : rt);
        }
        /* msquery.cpp:519 */
        orc_keyval_sort(key, val, (uint32_t)S);
        uint32_t speclen = 0;
        if (S > 0) {
            /* msquery.cpp:524-545 */
            const double factor = (double)base_int / key[S - 1];
            key[S - 1] = base_int;
            speclen = 1;
            for (int64_t j = S - 2; j >= S - (int64_t)qalen && j >= 0; j--) {
                double v = key[j] * factor;
                key[j] = v == v ? (uint32_t)v : 0u;   /* 0 * inf: x86-64 gives 0 */
                if (key[j] >= min_int) speclen++;
            }
        }
        /* msquery.cpp:548-562 */
        const uint32_t keep = speclen >= qalen ? qalen : speclen;
        const uint32_t off = idx[s];
        idx[s + 1] = off + keep;
        memcpy(moz + off, val + S - keep, (size_t)keep * 4);
        memcpy(intensity + off, key + S - keep, (size_t)keep * 4);
    }
    free(key);
    free(val);
    return HCX_SUCCESS;
}
