/*
 * hcx_oracle.h -- CPU restatement of the reference's one-rank search core.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the reported CPU baseline.  The product path (libhcx.so) never
 * links, loads or calls it.
 *
 * Every function restates the reference algorithm at the file:line cited
 * beside it.  The reference build needs CMake-generated headers
 * (config.hpp and common.hpp, generated from the templates in cmake/Templates), so it is not compiled
 * here; each function is pinned by the checks listed in its comment.
 */
#ifndef HCX_ORACLE_H
#define HCX_ORACLE_H
#include <stdint.h>
#include "../include/hcx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* utils.cpp:41-98 residue + static-mod tables (float). */
float orc_aa_mass(char c);
float orc_stat_mod(char c);

/* utils.h:115-131: val[0]=val[1]=0, val[i]=log10(i)+val[i-1].
 * Pinned: closed form lgamma(i+1)/ln(10) (tests/test_oracle_pins.py). */
void orc_lgfact(double *out, uint32_t n);

/* utils.cpp:290-303.  Pinned: published monoisotopic peptide masses. */
float orc_pep_mass(const char *seq, uint32_t len);
/* utils.cpp:337-362 (first mod only, Q4).  Pinned: unmod mass + mod mass. */
float orc_mod_mass(const char *seq, uint32_t len, uint32_t mod_num, const hcx_params *p);

/* utils.cpp:231-276 / 378-505.  Returns the precursor mass; fills speclen
 * = 2*maxz*(len-1) bins laid out [b z1..zZ][y z1..zZ].  Pinned: textbook
 * b/y ion m/z within the per-residue truncation bound, complementarity. */
float orc_gen_spectrum(const char *seq, uint32_t len, const hcx_params *p, uint32_t *out);
float orc_gen_mod_spectrum(const char *seq, uint32_t len, const hcx_pep_entry *e,
                           const hcx_params *p, uint32_t *out);

/* dslim.cpp:46-138, 216-347, 361-388, 452-482: bA = exclusive scan of the bin
 * histogram (NB+1 entries, NB = max_mass*scale); iA = ion ids sorted by bin,
 * ascending within a bin.  Pinned: prefix-sum / permutation invariants and a
 * brute-force sort on tiny inputs. */
int orc_index_build(const hcx_params *p, uint32_t peplen, const char *seqs,
                    const hcx_pep_entry *entries, uint32_t n,
                    uint32_t *iA, uint32_t *bA);
/* orc_index_build over nlen lengths, OpenMP over the lengths (CPU baseline) */
int orc_index_build_all(const hcx_params *p, uint32_t nlen, const uint32_t *peplen, const char *const *seqs,
                        const hcx_pep_entry *const *entries, const uint32_t *n, uint32_t *const *iA,
                        uint32_t *const *bA, int threads);

/* dslim_query.cpp:733-763 + expeRT.cpp:870-906: one shard's partRes and
 * u16 histogram slices of searched spectra (res, hist).  Pinned: hand-traced
 * slices and the 65500-level encoding with its int wrap
 * (tests/test_oracle_pins.py). */
int orc_part_export(const hcx_result *res, const uint32_t *hist, uint32_t hbins, uint32_t nspec, int32_t qid0,
                    uint32_t slot, hcx_part_res *packs, uint16_t *slices);
/* expeRT.cpp:720-852: the merged-histogram tail fit; returns the e-value.
 * Pinned: a histogram whose survival function is exactly geometric (closed
 * form line), the one-bin range, the accumulate-from-1 denominator. */
double orc_tailfit_merged(const double *pdata, uint32_t hbins, int32_t pN, int32_t hyp, uint32_t *flags);
/* dslim_score.cpp:244-345 + expeRT.cpp:909-936.  Pinned: hand-built
 * shard sets (tie to the first shard, skipped empty shards, cpsms sum,
 * decoding above 65500 candidates). */
int orc_combine(const hcx_params *p, uint32_t nshards, uint32_t nspec, uint32_t slot,
                const hcx_part_res *const *packs, const uint16_t *const *slices, hcx_fresult *out, int32_t *keys);

/* dslim_query.cpp:854-994.  Window [minl, maxl] over masses sorted
 * ascending; returns 1 if valid.  Pinned: brute-force linear scan. */
int orc_window(const float *mass, uint32_t n, float precursor, double dM,
               int32_t *minl, int32_t *maxl);

/* One per-length index as the query sees it. */
typedef struct orc_index {
    uint32_t peplen;
    uint32_t n;            /* entries */
    const float *mass;     /* n masses (sorted) */
    const uint32_t *iA;
    const uint32_t *bA;
} orc_index;

/* dslim_query.cpp:515-842 + expeRT.cpp:572-718, serial per spectrum,
 * OpenMP over spectra (threads <= 0: library default).  hist (optional) is
 * num_specs * (10*max_hyperscore+2) u32; stats optional. */
int orc_search(const hcx_params *p, const orc_index *idx, int nidx,
               const hcx_queries_view *q, hcx_result *res, uint32_t *hist,
               hcx_stats *stats, int threads);

/* Candidate dump: every inserted candidate (shpk >= min_shp, hs > 0) of
 * spectrum s, in the reference insertion order (length, psid). */
int orc_candidates(const hcx_params *p, const orc_index *idx, int nidx,
                   const hcx_queries_view *q, int32_t s,
                   hcx_cand *out, uint64_t cap, uint64_t *n);

/* Work counters only (window searches + bA probes + lower/upper bounds). */
int orc_count_work(const hcx_params *p, const orc_index *idx, int nidx,
                   const hcx_queries_view *q, hcx_stats *stats, int threads);

/* expeRT.cpp:572-718 on one histogram (hbins doubles of counts). */
int orc_tailfit(const double *survival, uint32_t hbins, uint32_t cpsms, int32_t maxhyp,
                int32_t *mu, int32_t *beta, int32_t *minhyp, int32_t *nexthyp,
                uint32_t *flags);

/* dslim.cpp:422-428: heap capacity 2^floor(log2(t)+0.999)-1 (t=0 -> 1). */
int orc_heap_capacity(uint32_t topmatches);

/* minheap.h:86-277 exposed for the heap pin tests: replays hs[] (in order)
 * into a heap of capacity cap and returns the getMax() index into hs[]. */
int orc_heap_replay(const float *hs, uint32_t n, int32_t cap, int32_t *argmax,
                    int32_t *final_size);

/* keyval.h:48-132 KeyVal_Parallel_Internal: Hoare quicksort of (key, val)
 * pairs by key, pivot key[(left+right)/2], recursing on [left, j] and
 * [i, right].  The task split of keyval.h:96-105 only chooses which thread
 * sorts a sub-range (the sub-ranges are disjoint), never the result.  Unstable:
 * equal keys end in the order this partition scheme leaves them.  Pinned:
 * hand-traced tie arrays, sortedness and the pair multiset on random arrays,
 * numpy's argsort on distinct keys (tests/test_oracle_pins.py). */
void orc_keyval_sort(uint32_t *key, uint32_t *val, uint32_t n);

/* msquery.cpp:509-569 MSQuery::pickpeaks, once per spectrum as extractbatch
 * calls it (msquery.cpp:282-285).  Raw peaks as readspectrum stores them
 * (msquery.cpp:403-404): mz = (uint)(m/z * scale), intensity =
 * (uint)(I * 1000).  Per spectrum: KeyVal sort by intensity; factor =
 * base_int / top; top -> base_int; the next up to qalen-1 peaks scaled
 * (uint)(I * factor) and counted when >= min_int; the last speclen peaks kept
 * in ascending-intensity order.  precurse = (float)prec_mz, charges =
 * MAX(1, z), rtimes = (float)MAX(0.0, rtime) (msquery.cpp:515-517).  A top
 * intensity of 0 makes the reference compute (uint)(0 * inf), undefined in C;
 * x86-64 gives 0 and so does this function.  Output arrays hold
 * sum(min(S, qalen)) peaks.  Pinned: hand-computed spectra (scaling,
 * truncation, the min_int cut, qalen, ties at the cut). */
int orc_pickpeaks(const uint32_t *raw_mz, const uint32_t *raw_in, const uint32_t *raw_idx,
                  const double *prec_mz, const int32_t *z, const double *rtime, int32_t n,
                  uint32_t qalen, uint32_t base_int, uint32_t min_int,
                  uint32_t *moz, uint32_t *intensity, uint32_t *idx,
                  float *precurse, int32_t *charges, float *rtimes);

/* ---- peptide database stage (hcx_oracle_lbe.cpp) ---- */

/* mods.cpp:290-315 count() with MODS_Initialize's parse (mods.cpp:64-103):
 * forms of one sequence, the unmodified one included.  Pinned: the number of
 * forms MODS_ModList emits plus one, counted by brute force over all site
 * subsets (tests/test_oracle_pins.py). */
int orc_lbe_count(const char *seq, uint32_t len, const char *modconditions, uint64_t *forms);
/* lbe_internal.cpp:397-418: upper case, keep lines whose float mass lies in
 * [min_mass, max_mass].  Pinned: published peptide masses at the bounds. */
int orc_lbe_filter(const char *lines, uint32_t nlines, uint32_t len, uint32_t min_mass, uint32_t max_mass,
                   char *seqs, uint32_t *nseqs);
/* LBE_Initialize (lbe_internal.cpp:119-202) of one length after
 * LBE_CountPeps: MODS_ModCounter, LBE_CreatePartitions, LBE_GeneratePeps,
 * MODS_GenerateMods (MODS_ModList order, glibc qsort with cmpvarEntries) and,
 * when sort != 0, std::sort by mass.  counts: pepCount, modCount, lclpepCnt,
 * lclmodCnt; out (lclpep + lclmod entries) may be NULL to get the counts.
 * Pinned: hand-traced variant lists and orders (tests/test_oracle_pins.py);
 * the sorts are the library calls the reference makes. */
int orc_lbe_generate(const hcx_params *p, const char *seqs, uint32_t nseqs, uint32_t len,
                     const char *modconditions, uint32_t nodes, uint32_t myid, int policy, int sort,
                     hcx_pep_entry *out, uint32_t *counts);
/* std::sort with CmpPepEntries on (key, value) records compared by key. */
void orc_std_sort_mass(float *keys, uint32_t *vals, uint32_t n);
/* glibc qsort with cmpvarEntries on (sites, value) records. */
void orc_qsort_sites(uint64_t *sites, uint32_t *vals, uint32_t n);
/* McIlroy's adversary against std::sort: n keys that drive it to its
 * heap-sort fallback. */
void orc_antiqsort(uint32_t n, float *keys);

#ifdef __cplusplus
}
#endif
#endif
