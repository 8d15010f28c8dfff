/*
 * hcx_datagen.h -- seeded synthetic peptide databases and MS2 query batches.
 *
 * Produces exactly what the search core's boundary consumes: per peptide
 * length, the sequence table and the mass-sorted pepEntries of the
 * reference's LBE stage (lbe_internal.cpp:119-229, mods.cpp:333-496), and a
 * CSR Queries batch normalised the way MSQuery::pickpeaks does
 * (msquery.cpp:509-569).  Counter-based RNG: the output depends only on the
 * configuration and the seed.
 */
#ifndef HCX_DATAGEN_H
#define HCX_DATAGEN_H
#include <stdint.h>
#include "hcx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hcx_dbgen_cfg {
    uint64_t seed;
    uint32_t npep;          /* base peptides (unmodified, mass-legal)           */
    uint32_t min_len;       /* lengths uniform in [min_len, max_len]            */
    uint32_t max_len;
    uint32_t min_mass;      /* legality filter on the unmodified mass            */
    uint32_t max_mass;
    uint32_t scale;
    /* variable mods: residues per mod type, per-type cap, per-peptide cap */
    uint32_t nmods;
    uint32_t max_mods;                       /* total per peptide (0: none)   */
    char     mod_residues[HCX_MAX_MOD_TYPES][4];
    uint32_t mod_max[HCX_MAX_MOD_TYPES];     /* per type per peptide           */
    uint32_t mod_mass[HCX_MAX_MOD_TYPES];    /* (uint)(Da*scale)                */
} hcx_dbgen_cfg;

typedef struct hcx_db hcx_db;

int  hcx_db_create(const hcx_dbgen_cfg *cfg, hcx_db **out);
int  hcx_db_free(hcx_db *db);
/* lengths with at least one entry, ascending */
int  hcx_db_lengths(const hcx_db *db, uint32_t *lens, uint32_t *nlens);
int  hcx_db_length_info(const hcx_db *db, uint32_t len, uint32_t *nseqs, uint32_t *nentries);
int  hcx_db_length_copy(const hcx_db *db, uint32_t len, char *seqs, hcx_pep_entry *entries);
/* zero-copy views, valid until hcx_db_free */
int  hcx_db_length_view(const hcx_db *db, uint32_t len, const char **seqs,
                        const hcx_pep_entry **entries, const float **mass);

typedef struct hcx_specgen_cfg {
    uint64_t seed;
    uint32_t nspec;
    uint32_t min_charge;     /* precursor charge uniform in [min, max]            */
    uint32_t max_charge;
    uint32_t frag_maxz;      /* fragment charges 1..frag_maxz                     */
    uint32_t n_noise;        /* noise peaks (at least the spectrum's peak count)  */
    uint32_t peaks_min;      /* kept peaks (QALEN): fixed if min == max,          */
    uint32_t peaks_max;      /*   else log-uniform in [min, max]                  */
    uint32_t scale;
    uint32_t base_int;       /* pickpeaks base intensity                          */
    uint32_t min_int;        /* pickpeaks floor                                    */
    uint32_t window_mode;    /* 0: uniform source; 1: source chosen for a window  */
    double   keep_prob;      /* b/y ion retention probability                      */
    double   noise_lo, noise_hi; /* noise m/z range                                */
    double   jitter_sd;      /* Gaussian m/z jitter (Da)                           */
    double   sig_mu, sig_sigma;     /* log-normal signal intensities                */
    double   noise_mu, noise_sigma; /* log-normal noise intensities                 */
    double   dM;             /* window_mode 1: the search's precursor tolerance    */
    double   wmin, wmax;     /* window_mode 1: target window sizes (log-uniform)  */
} hcx_specgen_cfg;

typedef struct hcx_specset hcx_specset;

int  hcx_spec_create(const hcx_db *db, const hcx_specgen_cfg *cfg, hcx_specset **out);
int  hcx_spec_free(hcx_specset *ss);
int  hcx_spec_info(const hcx_specset *ss, uint32_t *nspec, uint64_t *npeaks);
/* moz/intensity: npeaks; idx: nspec+1; precurse, charges, rtimes: nspec;
 * src_len / src_psid: the generating form (length, index in sorted entries). */
int  hcx_spec_copy(const hcx_specset *ss, uint32_t *moz, uint32_t *intensity, uint32_t *idx,
                   float *precurse, int32_t *charges, float *rtimes,
                   uint32_t *src_len, int32_t *src_psid);

/* The same spectra, also keeping each one's raw peaks before pickpeaks (in
 * ascending m/z, as an MS2 file lists them): mz = (uint)(m/z * scale),
 * intensity = (uint)(I * 1000) (msquery.cpp:403-404).  The normalised
 * spectra of hcx_spec_copy break intensity ties by a stable sort; the raw
 * peaks are for the device pickpeaks, which follows the reference's sort. */
int  hcx_spec_create_raw(const hcx_db *db, const hcx_specgen_cfg *cfg, hcx_specset **out);
int  hcx_spec_raw_info(const hcx_specset *ss, uint64_t *npeaks);
/* idx: nspec+1; prec_mz, z, rtime: nspec (the precursor mass, charge and
 * retention time of hcx_spec_copy, in double where the reader keeps double). */
int  hcx_spec_raw_copy(const hcx_specset *ss, uint32_t *mz, uint32_t *intensity, uint32_t *idx,
                       double *prec_mz, int32_t *z, double *rtime);

#ifdef __cplusplus
}
#endif
#endif
