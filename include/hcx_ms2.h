/*
 * hcx_ms2.h -- MS2 text ingest, the host side of the Queries producer.
 *
 * Reads a whole .ms2 file the way the reference's MSQuery does: the file
 * pre-scan of MSQuery::initialize (source/core/msquery.cpp:96-170) counts the
 * spectra ("S" lines), then MSQuery::readspectrum (msquery.cpp:300-505) is
 * applied once per spectrum.  The result is the raw input of the device
 * pickpeaks (hcx_raw_spectra_view in hcx.h):
 *   mz        = (uint)(atof(m/z token) * scale)      msquery.cpp:403
 *   intensity = (uint)(atof(intensity token) * 1000) msquery.cpp:404
 *   prec_mz   = atof of the Z line's third token     msquery.cpp:343
 *   z         = atoi of the Z line's second token
 *   rtime     = atof of an "I RTime <t>" line
 * with the reference's parsing rules kept as they are: Z / I lines split on
 * spaces and tabs, peak lines on spaces only; a missing peak token reads as
 * "0.01" in a file's first spectrum and "0.0" after it; Z, precursor and
 * retention time carry over to a spectrum that has no line for them (they
 * start at 0); peak lines before the first S belong to the first spectrum;
 * lines are split at '\n' only, so a '\r' stays part of the line.
 * Lines starting with H or D, and empty lines, are skipped.
 */
#ifndef HCX_MS2_H
#define HCX_MS2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hcx_ms2 hcx_ms2;

/* scale: bins per Da (gParams.scale).  HCX_ERR_FILE_NOT_FOUND if the
 * file cannot be opened, HCX_ERR_INVLD_SIZE if it holds no peak line (the
 * reference's initialize fails the same way). */
int hcx_ms2_read(const char *path, uint32_t scale, hcx_ms2 **out);
int hcx_ms2_free(hcx_ms2 *m);
/* nspec spectra, npeaks raw peaks in all, maxlen the largest spectrum */
int hcx_ms2_info(const hcx_ms2 *m, int32_t *nspec, uint64_t *npeaks, uint32_t *maxlen);
/* mz / intensity: npeaks; idx: nspec+1; prec_mz, z, rtime: nspec (any may be NULL) */
int hcx_ms2_copy(const hcx_ms2 *m, uint32_t *mz, uint32_t *intensity, uint32_t *idx,
                 double *prec_mz, int32_t *z, double *rtime);

#ifdef __cplusplus
}
#endif
#endif
