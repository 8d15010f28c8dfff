// datagen.cpp -- seeded synthetic peptide databases and MS2 query batches.
//
// The database side produces what the reference's LBE stage hands to
// DSLIM_Construct: for each peptide length a sequence table and the
// pepEntries sorted by precursor mass (lbe_internal.cpp:119-229; mod variants
// from mods.cpp:333-496, masses from utils.cpp:290-362 including the
// first-mod-only precursor mass).  Ties in the mass sort keep generation order
// (stable sort); the reference's std::sort leaves tie order unspecified, so
// the order here is this generator's definition and both the oracle and the
// GPU consume the same arrays.
//
// The query side builds spectra the way MSQuery::readspectrum + pickpeaks
// (msquery.cpp:300-569) would store them: m/z -> (uint)(mz*scale), intensity
// -> (uint)(I*1000), sorted ascending by intensity, the top peak scaled to
// base_int, the rest scaled by base_int/max and truncated, top-QALEN peaks at
// or above min_int kept, stored in ascending-intensity order.
#include "../../include/hcx_datagen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace {

constexpr float kH2O = 18.015f;       // utils.cpp:29
constexpr double kProton = 1.00727647; // utils.cpp:30
constexpr float kNAA = -20000.0f;

// utils.cpp:41-98: residue masses (float) and the static C modification.
const float kAAMass[26] = {
    71.03712f, kNAA, 103.00919f, 115.030f, 129.0426f, 147.068f, 57.02146f,
    137.060f, 113.084f, kNAA, 128.094f, 113.084f, 131.0405f, 114.043f, kNAA,
    97.0527f, 128.05858f, 156.1012f, 87.032f, 101.0476f, kNAA, 99.06841f,
    186.0793f, kNAA, 163.0633f, kNAA};
inline float stat_mod(char c) { return c == 'C' ? 57.021464f : 0.0f; }

const char kResidues[] = "ACDEFGHIKLMNPQRSTVWY";

inline uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) : s(seed ^ (stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull)) {
        splitmix(s);
    }
    uint64_t u64() { return splitmix(s); }
    double unif() { return (double)(u64() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t below(uint32_t n) { return (uint32_t)((u64() >> 32) * (uint64_t)n >> 32); }
    double normal() {
        double u1 = unif(), u2 = unif();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// utils.cpp:290-303
float pep_mass(const char *seq, uint32_t len) {
    float m = kH2O;
    for (uint32_t l = 0; l < len; l++) {
        m += kAAMass[seq[l] - 'A'];
        m += stat_mod(seq[l]);
    }
    return m;
}

// utils.cpp:337-362: only the first mod nibble reaches the mass (the loop's
// second shift count is the mask itself, which x86 reduces to 16 bits).
float mod_mass(const char *seq, uint32_t len, uint32_t mod_num, const uint32_t *mm, uint32_t scale) {
    float m = pep_mass(seq, len);
    uint32_t first = mod_num & 0xF;
    if (first) m += (float)mm[first - 1] / (float)scale;
    return m;
}

struct LenTable {
    std::vector<char> seqs;
    std::vector<hcx_pep_entry> entries;
    std::vector<float> mass;
    uint32_t nseqs = 0;
};

}  // namespace

struct hcx_db {
    hcx_dbgen_cfg cfg;
    uint32_t min_len = 0, max_len = 0;
    std::vector<LenTable> lens;   // index len - min_len
    const LenTable *get(uint32_t L) const {
        if (L < min_len || L > max_len) return nullptr;
        return &lens[L - min_len];
    }
};

struct hcx_specset {
    std::vector<uint32_t> moz, inten, idx, src_len;
    std::vector<float> prec, rt;
    std::vector<int32_t> chg, src_psid;
    // raw peaks before pickpeaks (hcx_spec_create_raw only)
    std::vector<uint32_t> raw_mz, raw_in, raw_idx;
};

namespace {

// mods.cpp:333-384: recursive variant enumeration (take the mod first, then
// skip it); modNum gets (type + 1) in the nibble of the mod's rank.
void enum_mods(const char *seq, uint32_t len, const int *type_of, int32_t *cond_left, uint32_t total,
               uint32_t letter, bool novel, uint32_t seen, uint64_t sites, uint32_t modnum,
               std::vector<std::pair<uint64_t, uint32_t>> &out) {
    if (novel && letter != 0) out.emplace_back(sites, modnum);
    if (total == 0 || letter >= len) return;
    int t = type_of[seq[letter] - 'A'];
    if (t != -1 && cond_left[t] > 0) {
        cond_left[t] -= 1;
        enum_mods(seq, len, type_of, cond_left, total - 1, letter + 1, true, seen + 1,
                  sites | (1ull << letter), modnum + (1u << (4 * seen)) * (uint32_t)(1 + t), out);
        cond_left[t] += 1;
    }
    enum_mods(seq, len, type_of, cond_left, total, letter + 1, false, seen, sites, modnum, out);
}

}  // namespace

extern "C" int hcx_db_create(const hcx_dbgen_cfg *cfg, hcx_db **out) {
    if (!cfg || !out) return HCX_ERR_INVLD_PTR;
    if (cfg->min_len < 2 || cfg->max_len < cfg->min_len || cfg->max_len > 60 || cfg->scale == 0)
        return HCX_ERR_INVLD_PARAM;
    if (cfg->nmods > HCX_MAX_MOD_TYPES) return HCX_ERR_INVLD_PARAM;
    hcx_db *db = new (std::nothrow) hcx_db;
    if (!db) return HCX_ERR_BAD_MEM_ALLOC;
    db->cfg = *cfg;
    db->min_len = cfg->min_len;
    db->max_len = cfg->max_len;
    db->lens.resize(cfg->max_len - cfg->min_len + 1);
    const uint32_t span = cfg->max_len - cfg->min_len + 1;

    // 1. candidate peptides in counter order; keep the first npep mass-legal
    //    ones (LBE_CountPeps mass filter, lbe_internal.cpp:411-418).
    std::vector<uint32_t> acc_len;
    std::vector<char> acc_seq;   // max_len chars per accepted peptide
    acc_len.reserve(cfg->npep);
    acc_seq.reserve((size_t)cfg->npep * cfg->max_len);
    const uint64_t chunk = 1u << 20;
    std::vector<char> cseq(chunk * cfg->max_len);
    std::vector<uint32_t> clen(chunk);
    std::vector<uint8_t> ok(chunk);
    for (uint64_t base = 0; acc_len.size() < cfg->npep; base += chunk) {
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < (int64_t)chunk; j++) {
            Rng r(cfg->seed, base + (uint64_t)j);
            uint32_t L = cfg->min_len + r.below(span);
            char *s = &cseq[(size_t)j * cfg->max_len];
            for (uint32_t l = 0; l + 1 < L; l++) s[l] = kResidues[r.below(20)];
            s[L - 1] = (r.u64() >> 63) ? 'K' : 'R';
            float m = pep_mass(s, L);
            clen[j] = L;
            ok[j] = (m >= (float)cfg->min_mass && m <= (float)cfg->max_mass) ? 1 : 0;
        }
        for (uint64_t j = 0; j < chunk && acc_len.size() < cfg->npep; j++) {
            if (!ok[j]) continue;
            acc_len.push_back(clen[j]);
            acc_seq.insert(acc_seq.end(), &cseq[j * cfg->max_len], &cseq[j * cfg->max_len] + cfg->max_len);
        }
        if (base > (uint64_t)cfg->npep * 64 + (1u << 24)) { delete db; return HCX_ERR_INVLD_PARAM; }
    }

    // 2. group by length (acceptance order = seqID order within a length)
    for (size_t i = 0; i < acc_len.size(); i++) {
        LenTable &T = db->lens[acc_len[i] - cfg->min_len];
        T.seqs.insert(T.seqs.end(), &acc_seq[i * cfg->max_len], &acc_seq[i * cfg->max_len] + acc_len[i]);
        T.nseqs++;
    }

    int type_of[26];
    for (int i = 0; i < 26; i++) type_of[i] = -1;
    for (uint32_t t = 0; t < cfg->nmods; t++)
        for (int c = 0; c < 4 && cfg->mod_residues[t][c]; c++) {
            char ch = cfg->mod_residues[t][c];
            if (ch >= 'A' && ch <= 'Z') type_of[ch - 'A'] = (int)t;
        }

    // 3. per length: unmodified entries, then variants, then a stable mass sort
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t li = 0; li < (int64_t)span; li++) {
        LenTable &T = db->lens[li];
        const uint32_t L = cfg->min_len + (uint32_t)li;
        std::vector<hcx_pep_entry> ents;
        ents.reserve(T.nseqs);
        for (uint32_t i = 0; i < T.nseqs; i++) {
            hcx_pep_entry e;
            std::memset(&e, 0, sizeof e);
            e.mass = pep_mass(&T.seqs[(size_t)i * L], L);
            e.seq_id = i;
            ents.push_back(e);
        }
        if (cfg->max_mods > 0 && cfg->nmods > 0) {
            std::vector<std::pair<uint64_t, uint32_t>> vars;
            int32_t cond[HCX_MAX_MOD_TYPES];
            for (uint32_t i = 0; i < T.nseqs; i++) {
                const char *s = &T.seqs[(size_t)i * L];
                for (uint32_t t = 0; t < cfg->nmods; t++) cond[t] = (int32_t)cfg->mod_max[t];
                vars.clear();
                enum_mods(s, L, type_of, cond, cfg->max_mods, 0, false, 0, 0, 0, vars);
                for (auto &v : vars) {
                    hcx_pep_entry e;
                    std::memset(&e, 0, sizeof e);
                    e.seq_id = i;
                    e.sites = v.first;
                    e.mod_num = v.second;
                    e.mass = mod_mass(s, L, v.second, cfg->mod_mass, cfg->scale);
                    ents.push_back(e);
                }
            }
        }
        std::stable_sort(ents.begin(), ents.end(),
                         [](const hcx_pep_entry &a, const hcx_pep_entry &b) { return a.mass < b.mass; });
        T.entries.swap(ents);
        T.mass.resize(T.entries.size());
        for (size_t k = 0; k < T.entries.size(); k++) T.mass[k] = T.entries[k].mass;
    }
    *out = db;
    return HCX_SUCCESS;
}

extern "C" int hcx_db_free(hcx_db *db) {
    delete db;
    return HCX_SUCCESS;
}

extern "C" int hcx_db_lengths(const hcx_db *db, uint32_t *lens, uint32_t *nlens) {
    if (!db || !nlens) return HCX_ERR_INVLD_PTR;
    uint32_t n = 0;
    for (uint32_t L = db->min_len; L <= db->max_len; L++)
        if (!db->get(L)->entries.empty()) {
            if (lens) lens[n] = L;
            n++;
        }
    *nlens = n;
    return HCX_SUCCESS;
}

extern "C" int hcx_db_length_info(const hcx_db *db, uint32_t len, uint32_t *nseqs, uint32_t *nentries) {
    if (!db) return HCX_ERR_INVLD_PTR;
    const LenTable *T = db->get(len);
    if (!T) return HCX_ERR_INVLD_PARAM;
    if (nseqs) *nseqs = T->nseqs;
    if (nentries) *nentries = (uint32_t)T->entries.size();
    return HCX_SUCCESS;
}

extern "C" int hcx_db_length_copy(const hcx_db *db, uint32_t len, char *seqs, hcx_pep_entry *entries) {
    if (!db) return HCX_ERR_INVLD_PTR;
    const LenTable *T = db->get(len);
    if (!T) return HCX_ERR_INVLD_PARAM;
    if (seqs && !T->seqs.empty()) std::memcpy(seqs, T->seqs.data(), T->seqs.size());
    if (entries && !T->entries.empty())
        std::memcpy(entries, T->entries.data(), T->entries.size() * sizeof(hcx_pep_entry));
    return HCX_SUCCESS;
}

extern "C" int hcx_db_length_view(const hcx_db *db, uint32_t len, const char **seqs,
                                  const hcx_pep_entry **entries, const float **mass) {
    if (!db) return HCX_ERR_INVLD_PTR;
    const LenTable *T = db->get(len);
    if (!T) return HCX_ERR_INVLD_PARAM;
    if (seqs) *seqs = T->seqs.data();
    if (entries) *entries = T->entries.data();
    if (mass) *mass = T->mass.data();
    return HCX_SUCCESS;
}

namespace {

struct FormRef { float mass; uint32_t len; int32_t psid; };

struct Peak { uint32_t mz; uint32_t in; };

// Fragment m/z of one b or y ion: neutral residue sum (+ mods, + H2O for y),
// plus z protons, over z.
void spectrum_for(const hcx_db *db, const hcx_specgen_cfg *cfg, uint32_t L, int32_t psid, uint32_t s,
                  std::vector<Peak> &peaks_out, uint32_t qalen, std::vector<Peak> *raw_out) {
    const LenTable *T = db->get(L);
    const hcx_pep_entry &e = T->entries[psid];
    const char *seq = &T->seqs[(size_t)e.seq_id * L];
    Rng r(cfg->seed ^ 0x5bd1e995ull, 0x100000000ull + s);
    double res[64], add[64];
    int types[16];
    int nt = 0;
    for (int k = 0; k < 15; k++) {
        uint32_t nib = (e.mod_num >> (4 * k)) & 0xF;
        if (nib) types[nt++] = (int)nib - 1;
    }
    int seen = 0;
    for (uint32_t l = 0; l < L; l++) {
        res[l] = (double)(kAAMass[seq[l] - 'A'] + stat_mod(seq[l]));
        add[l] = 0.0;
        if ((e.sites >> l) & 1ull) {
            int t = seen < nt ? types[seen] : 0;
            add[l] = (double)db->cfg.mod_mass[t] / (double)db->cfg.scale;
            seen++;
        }
    }
    std::vector<std::pair<double, double>> raw;   // (mz, intensity)
    for (uint32_t z = 1; z <= cfg->frag_maxz; z++) {
        double bsum = 0.0, ysum = (double)kH2O;
        for (uint32_t i = 1; i < L; i++) {
            bsum += res[i - 1] + add[i - 1];
            ysum += res[L - i] + add[L - i];
            for (int series = 0; series < 2; series++) {
                double m = series == 0 ? bsum : ysum;
                if (r.unif() >= cfg->keep_prob) continue;
                double mz = (m + z * kProton) / z + cfg->jitter_sd * r.normal();
                double in = std::exp(cfg->sig_mu + cfg->sig_sigma * r.normal());
                if (mz > 0) raw.emplace_back(mz, in);
            }
        }
    }
    uint32_t nn = std::max(cfg->n_noise, qalen);
    for (uint32_t i = 0; i < nn; i++) {
        double mz = cfg->noise_lo + (cfg->noise_hi - cfg->noise_lo) * r.unif();
        double in = std::exp(cfg->noise_mu + cfg->noise_sigma * r.normal());
        raw.emplace_back(mz, in);
    }
    // readspectrum (msquery.cpp:403-404) then pickpeaks (msquery.cpp:509-569)
    std::vector<Peak> pk(raw.size());
    for (size_t i = 0; i < raw.size(); i++) {
        pk[i].mz = (uint32_t)(raw[i].first * cfg->scale);
        pk[i].in = (uint32_t)(raw[i].second * 1000);
    }
    if (raw_out) {
        // as an MS2 file lists them: ascending m/z
        *raw_out = pk;
        std::stable_sort(raw_out->begin(), raw_out->end(), [](const Peak &a, const Peak &b) { return a.mz < b.mz; });
    }
    std::stable_sort(pk.begin(), pk.end(), [](const Peak &a, const Peak &b) { return a.in < b.in; });
    const int S = (int)pk.size();
    uint32_t speclen = 0;
    if (S > 0 && pk[S - 1].in > 0) {
        double factor = (double)cfg->base_int / pk[S - 1].in;
        pk[S - 1].in = cfg->base_int;
        speclen = 1;
        for (int j = S - 2; j >= S - (int)qalen && j >= 0; j--) {
            pk[j].in = (uint32_t)(pk[j].in * factor);
            if (pk[j].in >= cfg->min_int) speclen++;
        }
    }
    if (speclen > qalen) speclen = qalen;
    peaks_out.assign(pk.end() - speclen, pk.end());
}

}  // namespace

namespace {

int spec_create(const hcx_db *db, const hcx_specgen_cfg *cfg, bool keep_raw, hcx_specset **out) {
    if (!db || !cfg || !out) return HCX_ERR_INVLD_PTR;
    if (cfg->peaks_min == 0 || cfg->peaks_max < cfg->peaks_min || cfg->frag_maxz == 0)
        return HCX_ERR_INVLD_PARAM;
    std::vector<FormRef> forms;
    for (uint32_t L = db->min_len; L <= db->max_len; L++) {
        const LenTable *T = db->get(L);
        for (size_t k = 0; k < T->entries.size(); k++)
            forms.push_back({T->entries[k].mass, L, (int32_t)k});
    }
    if (forms.empty()) return HCX_ERR_INVLD_PARAM;
    std::stable_sort(forms.begin(), forms.end(), [](const FormRef &a, const FormRef &b) { return a.mass < b.mass; });
    const size_t NF = forms.size();

    // window_mode 1: bucket forms by log10(#forms within +-dM of their mass)
    std::vector<std::vector<uint32_t>> bucket;
    const int NBK = 80;
    if (cfg->window_mode == 1) {
        bucket.resize(NBK);
        size_t lo = 0, hi = 0;
        for (size_t c = 0; c < NF; c++) {
            double m = forms[c].mass;
            while (lo < NF && forms[lo].mass < m - cfg->dM) lo++;
            while (hi < NF && forms[hi].mass <= m + cfg->dM) hi++;
            double w = (double)(hi - lo);
            int b = (int)std::floor(std::log10(std::max(w, 1.0)) * 10.0);
            if (b >= NBK) b = NBK - 1;
            bucket[b].push_back((uint32_t)c);
        }
    }

    hcx_specset *ss = new (std::nothrow) hcx_specset;
    if (!ss) return HCX_ERR_BAD_MEM_ALLOC;
    const uint32_t n = cfg->nspec;
    std::vector<std::vector<Peak>> per(n), per_raw(keep_raw ? n : 0);
    ss->prec.resize(n);
    ss->chg.resize(n);
    ss->rt.resize(n);
    ss->src_len.resize(n);
    ss->src_psid.resize(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < (int64_t)n; s++) {
        Rng r(cfg->seed, (uint64_t)s);
        size_t c;
        if (cfg->window_mode == 1) {
            double lw = std::log10(cfg->wmin) + (std::log10(cfg->wmax) - std::log10(cfg->wmin)) * r.unif();
            int b = (int)std::floor(lw * 10.0);
            if (b < 0) b = 0;
            if (b >= NBK) b = NBK - 1;
            int bb = -1;
            for (int d = 0; d < NBK && bb < 0; d++) {
                if (b - d >= 0 && !bucket[b - d].empty()) bb = b - d;
                else if (b + d < NBK && !bucket[b + d].empty()) bb = b + d;
            }
            const auto &B = bucket[bb];
            c = B[r.below((uint32_t)B.size())];
        } else {
            c = (size_t)(r.unif() * NF);
            if (c >= NF) c = NF - 1;
        }
        uint32_t qalen = cfg->peaks_min;
        if (cfg->peaks_max > cfg->peaks_min) {
            double lq = std::log((double)cfg->peaks_min) +
                        (std::log((double)cfg->peaks_max) - std::log((double)cfg->peaks_min)) * r.unif();
            qalen = (uint32_t)std::lround(std::exp(lq));
            qalen = std::min(std::max(qalen, cfg->peaks_min), cfg->peaks_max);
        }
        ss->prec[s] = forms[c].mass;
        ss->chg[s] = (int32_t)(cfg->min_charge + r.below(cfg->max_charge - cfg->min_charge + 1));
        ss->rt[s] = (float)s * 0.01f;
        ss->src_len[s] = forms[c].len;
        ss->src_psid[s] = forms[c].psid;
        spectrum_for(db, cfg, forms[c].len, forms[c].psid, (uint32_t)s, per[s], qalen,
                     keep_raw ? &per_raw[s] : nullptr);
    }
    ss->idx.resize(n + 1);
    ss->idx[0] = 0;
    for (uint32_t s = 0; s < n; s++) ss->idx[s + 1] = ss->idx[s] + (uint32_t)per[s].size();
    ss->moz.resize(ss->idx[n]);
    ss->inten.resize(ss->idx[n]);
    for (uint32_t s = 0; s < n; s++)
        for (size_t k = 0; k < per[s].size(); k++) {
            ss->moz[ss->idx[s] + k] = per[s][k].mz;
            ss->inten[ss->idx[s] + k] = per[s][k].in;
        }
    if (keep_raw) {
        ss->raw_idx.resize(n + 1);
        ss->raw_idx[0] = 0;
        for (uint32_t s = 0; s < n; s++) ss->raw_idx[s + 1] = ss->raw_idx[s] + (uint32_t)per_raw[s].size();
        ss->raw_mz.resize(ss->raw_idx[n]);
        ss->raw_in.resize(ss->raw_idx[n]);
        for (uint32_t s = 0; s < n; s++)
            for (size_t k = 0; k < per_raw[s].size(); k++) {
                ss->raw_mz[ss->raw_idx[s] + k] = per_raw[s][k].mz;
                ss->raw_in[ss->raw_idx[s] + k] = per_raw[s][k].in;
            }
    }
    *out = ss;
    return HCX_SUCCESS;
}

}  // namespace

extern "C" int hcx_spec_create(const hcx_db *db, const hcx_specgen_cfg *cfg, hcx_specset **out) {
    return spec_create(db, cfg, false, out);
}

extern "C" int hcx_spec_create_raw(const hcx_db *db, const hcx_specgen_cfg *cfg, hcx_specset **out) {
    return spec_create(db, cfg, true, out);
}

extern "C" int hcx_spec_raw_info(const hcx_specset *ss, uint64_t *npeaks) {
    if (!ss) return HCX_ERR_INVLD_PTR;
    if (npeaks) *npeaks = ss->raw_mz.size();
    return ss->raw_idx.empty() ? HCX_ERR_INVLD_PARAM : HCX_SUCCESS;
}

extern "C" int hcx_spec_raw_copy(const hcx_specset *ss, uint32_t *mz, uint32_t *intensity, uint32_t *idx,
                                 double *prec_mz, int32_t *z, double *rtime) {
    if (!ss) return HCX_ERR_INVLD_PTR;
    if (ss->raw_idx.empty()) return HCX_ERR_INVLD_PARAM;
    const size_t n = ss->prec.size();
    if (mz) std::memcpy(mz, ss->raw_mz.data(), ss->raw_mz.size() * 4);
    if (intensity) std::memcpy(intensity, ss->raw_in.data(), ss->raw_in.size() * 4);
    if (idx) std::memcpy(idx, ss->raw_idx.data(), ss->raw_idx.size() * 4);
    for (size_t s = 0; s < n; s++) {
        if (prec_mz) prec_mz[s] = (double)ss->prec[s];
        if (z) z[s] = ss->chg[s];
        if (rtime) rtime[s] = (double)ss->rt[s];
    }
    return HCX_SUCCESS;
}

extern "C" int hcx_spec_free(hcx_specset *ss) {
    delete ss;
    return HCX_SUCCESS;
}

extern "C" int hcx_spec_info(const hcx_specset *ss, uint32_t *nspec, uint64_t *npeaks) {
    if (!ss) return HCX_ERR_INVLD_PTR;
    if (nspec) *nspec = (uint32_t)ss->prec.size();
    if (npeaks) *npeaks = ss->moz.size();
    return HCX_SUCCESS;
}

extern "C" int hcx_spec_copy(const hcx_specset *ss, uint32_t *moz, uint32_t *intensity, uint32_t *idx,
                             float *precurse, int32_t *charges, float *rtimes, uint32_t *src_len,
                             int32_t *src_psid) {
    if (!ss) return HCX_ERR_INVLD_PTR;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(moz, ss->moz.data(), ss->moz.size() * 4);
    cp(intensity, ss->inten.data(), ss->inten.size() * 4);
    cp(idx, ss->idx.data(), ss->idx.size() * 4);
    cp(precurse, ss->prec.data(), ss->prec.size() * 4);
    cp(charges, ss->chg.data(), ss->chg.size() * 4);
    cp(rtimes, ss->rt.data(), ss->rt.size() * 4);
    cp(src_len, ss->src_len.data(), ss->src_len.size() * 4);
    cp(src_psid, ss->src_psid.data(), ss->src_psid.size() * 4);
    return HCX_SUCCESS;
}
