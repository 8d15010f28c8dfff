// hcx_oracle_lbe.cpp -- CPU restatement of the reference's peptide database
// stage (LBE + MODS): the mass filter of LBE_CountPeps, the variant counts of
// MODS_ModCounter, the rank share of LBE_CreatePartitions, the entries of
// LBE_GeneratePeps and MODS_GenerateMods, and the mass sort of LBE_Initialize.
//
// TEST INFRASTRUCTURE ONLY (see hcx_oracle.h): loaded by tests/ and bench's
// cpu_baseline leg as the checker, never by the product path.
//
// The two sorts are the library calls the reference makes, with the
// reference's comparators restated: glibc qsort with cmpvarEntries
// (mods.cpp:105-126, 491; slm_dsts.h:191-271) and libstdc++ std::sort with
// CmpPepEntries (lbe_internal.cpp:184-187, 491; slm_dsts.h:176-179).  Their
// tie orders are those of this image's glibc 2.35 (qsort = a stable merge
// sort for arrays that fit its temporary buffer) and GCC 11 libstdc++
// (introsort), the toolchain the reference would be built with here.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "hcx_oracle.h"

namespace {

constexpr int kMaxSeqLen = 60;   // CMakeLists.txt:156-157 (MAX_SEQ_LEN)
constexpr int kMaxCombs = 64;    // mods.h:30 (MAX_COMBS)

// mods.cpp:137-160 MODS_GenCombinations: Pascal's triangle, filled the
// reference's way (left half by the recurrence, right half mirrored).
struct Comb {
    uint64_t c[kMaxCombs][kMaxCombs];
    Comb() {
        std::memset(c, 0, sizeof c);
        c[0][0] = 1;
        c[1][0] = 1;
        c[1][1] = 1;
        for (int i = 1; i < kMaxCombs; i++) {
            c[i][0] = 1;
            for (int j = 1; j <= (i + 1) / 2; j++) c[i][j] = c[i - 1][j - 1] + c[i - 1][j];
            for (int j = i / 2; j < i; j++) c[i][j] = c[i][i - j];
            c[i][i] = 1;
        }
    }
};
const Comb &comb() {
    static const Comb C;
    return C;
}

// mods.cpp:64-103 MODS_Initialize: "limit AAs n AAs n ..." split on white
// space; a residue's type is the last condition naming it.
struct Mods {
    int limit = 0;
    std::vector<std::string> residues;
    std::vector<int> cond;
    int lookup[256];
};

int parse_mods(const char *text, Mods *m) {
    std::stringstream ss(text ? text : "");
    std::vector<std::string> tok;
    std::string t;
    while (ss >> t) tok.push_back(t);
    if (tok.empty()) return HCX_ERR_INVLD_MOD;
    m->limit = std::atoi(tok[0].c_str());
    for (int i = 0; i < 256; i++) m->lookup[i] = -1;
    for (size_t i = 0; i < (tok.size() - 1) / 2; i++) {
        for (char c : tok[2 * i + 1]) m->lookup[(unsigned char)c] = (int)i;
        m->residues.push_back(tok[2 * i + 1]);
        m->cond.push_back(std::atoi(tok[2 * i + 2].c_str()));
    }
    return HCX_SUCCESS;
}

// mods.cpp:192-231 partition2: ways to pick up to b elements from the
// multiset whose class sizes are a.
int64_t partition2(const std::vector<int> &a, int b) {
    int64_t sum = 0;
    for (int t : a) sum += t;
    if (b > sum) b = (int)sum;
    if (b <= 0) return (int64_t)(b == 0);
    if (a.empty()) return 1;
    std::vector<int> a2(a.begin() + 1, a.end());
    sum = 0;
    for (int i = 0; i < a[0] + 1; i++) sum += (int64_t)comb().c[a[0]][i] * partition2(a2, b - i);
    return sum;
}

// mods.cpp:246-276 partition3: the per-condition caps and the per-peptide limit.
int64_t partition3(const std::vector<std::vector<int>> &A, const std::vector<int> &B, int limit) {
    if (A.size() == 1) return B[0] <= limit ? partition2(A[0], B[0]) : partition2(A[0], limit);
    std::vector<std::vector<int>> A2(A.begin() + 1, A.end());
    std::vector<int> B2(B.begin() + 1, B.end());
    int64_t sum = 0;
    for (int i = 0; i < B[0] + 1; i++)
        sum += (partition2(A[0], i) - partition2(A[0], i - 1)) * partition3(A2, B2, limit - i);
    return sum;
}

// mods.cpp:290-315 count(): forms of one sequence, the unmodified one included.
uint64_t count_forms(const char *s, uint32_t L, const Mods &m) {
    int cnt[256] = {0};
    for (uint32_t i = 0; i < L; i++) cnt[(unsigned char)s[i]]++;
    std::vector<std::vector<int>> A;
    std::vector<int> B;
    for (size_t i = 0; i < m.residues.size(); i++) {
        std::vector<int> t;
        for (char c : m.residues[i]) t.push_back(cnt[(unsigned char)c]);
        A.push_back(t);
        B.push_back(m.cond[i]);
    }
    if (A.empty()) return 1;
    return (uint64_t)partition3(A, B, m.limit);
}

// mods.cpp:333-384 MODS_ModList: depth-first, the branch that modifies the
// current residue before the one that skips it; a form is emitted on entry
// to every modifying branch.  keep(global) is LBE_ApplyPolicy
// (lbe_internal.cpp:75-103) for modified forms.
template <class Keep>
void mod_list(const char *pep, uint32_t L, std::vector<int> &cond, int total, hcx_pep_entry cont, uint32_t letter,
              bool novel, int seen, uint32_t refid, uint32_t &global, std::vector<hcx_pep_entry> &out, const Mods &m,
              const Keep &keep) {
    if (novel && letter != 0) {
        if (keep(global)) out.push_back(cont);
        global++;
    }
    if (total == 0 || letter >= L) return;
    int t = m.lookup[(unsigned char)pep[letter]];
    if (t != -1 && cond[t] > 0) {
        hcx_pep_entry d = cont;
        d.sites |= (uint64_t)1 << letter;
        d.mod_num += (1u << (4 * seen)) * (uint32_t)(1 + t);
        d.seq_id = refid;
        cond[t] -= 1;
        mod_list(pep, L, cond, total - 1, d, letter + 1, true, seen + 1, refid, global, out, m, keep);
        cond[t] += 1;
    }
    mod_list(pep, L, cond, total, cont, letter + 1, false, seen, refid, global, out, m, keep);
}

// slm_dsts.h:191-271: a form's site distance, s1 + MAX_SEQ_LEN - s3, from the
// first modified residue s1 and the last s3, found by the reference's loops.
int site_distance(uint64_t sites) {
    int s1 = 0, s3 = kMaxSeqLen;
    while (((sites >> s1) & 1) != 1 && s1++ < kMaxSeqLen) {
    }
    while (((sites >> s3) & 1) != 1 && s3-- > s1) {
    }
    return s1 + kMaxSeqLen - s3;
}

// mods.cpp:105-126 cmpvarEntries (operator<< / operator>>).
int cmp_var_entries(const void *l, const void *r) {
    const hcx_pep_entry *a = (const hcx_pep_entry *)l, *b = (const hcx_pep_entry *)r;
    int da = site_distance(a->sites), db = site_distance(b->sites);
    if (da < db) return 1;
    if (da > db) return -1;
    return 0;
}

// lbe_internal.cpp:491 CmpPepEntries = pepEntry::operator< (slm_dsts.h:176-179)
bool cmp_pep_entries(hcx_pep_entry a, hcx_pep_entry b) { return a.mass < b.mass; }

// dslim.cpp:766-780 DSLIM_GenerateIndex for the unmodified forms
uint32_t generate_index(uint32_t key, uint32_t lclpep, uint32_t nodes, uint32_t myid, int policy) {
    return policy == HCX_POLICY_CHUNK ? myid * lclpep + key : key * nodes + myid;
}

}  // namespace

extern "C" int orc_lbe_count(const char *seq, uint32_t len, const char *modconditions, uint64_t *forms) {
    Mods m;
    int rc = parse_mods(modconditions, &m);
    if (rc) return rc;
    *forms = count_forms(seq, len, m);
    return HCX_SUCCESS;
}

extern "C" int orc_lbe_filter(const char *lines, uint32_t nlines, uint32_t len, uint32_t min_mass, uint32_t max_mass,
                              char *seqs, uint32_t *nseqs) {
    // lbe_internal.cpp:397-418: upper case, then the float mass against the
    // integer bounds (converted to float by the comparison)
    uint32_t n = 0;
    std::vector<char> s(len);
    for (uint32_t i = 0; i < nlines; i++) {
        for (uint32_t l = 0; l < len; l++) {
            char c = lines[(size_t)i * len + l];
            s[l] = (c >= 'a' && c <= 'z') ? (char)(c - 32) : c;
        }
        float pm = orc_pep_mass(s.data(), len);
        if (pm >= (float)min_mass && pm <= (float)max_mass) {
            if (seqs) std::memcpy(seqs + (size_t)n * len, s.data(), len);
            n++;
        }
    }
    *nseqs = n;
    return HCX_SUCCESS;
}

extern "C" int orc_lbe_generate(const hcx_params *p, const char *seqs, uint32_t nseqs, uint32_t len,
                                const char *modconditions, uint32_t nodes, uint32_t myid, int policy, int sort,
                                hcx_pep_entry *out, uint32_t *counts) {
    if (!p || !counts || (nseqs && !seqs)) return HCX_ERR_INVLD_PTR;
    if (nodes == 0 || myid >= nodes) return HCX_ERR_INVLD_NODE_RANK;
    Mods m;
    int rc = parse_mods(modconditions, &m);
    if (rc) return rc;
    // mods.cpp:399-432 MODS_ModCounter: forms - 1 per sequence, exclusive scan
    std::vector<uint32_t> var_count(nseqs + 1, 0);
    uint64_t mod_count = 0;
    if (m.limit > 0) {
        for (uint32_t i = 0; i < nseqs; i++) {
            uint64_t v = count_forms(seqs + (size_t)i * len, len, m) - 1;
            var_count[i + 1] = var_count[i] + (uint32_t)v;
            mod_count += v;
        }
    }
    // lbe_internal.cpp:308-352 LBE_CreatePartitions
    uint32_t lclpep = nseqs, lclmod = (uint32_t)mod_count;
    if (nodes > 1) {
        lclpep = nseqs / nodes + ((nseqs % nodes) > myid ? 1 : 0);
        lclmod = (uint32_t)(mod_count / nodes + ((mod_count % nodes) > myid ? 1 : 0));
    }
    counts[0] = nseqs;
    counts[1] = (uint32_t)mod_count;
    counts[2] = lclpep;
    counts[3] = lclmod;
    if (!out) return HCX_SUCCESS;
    // lbe_internal.cpp:204-229 LBE_GeneratePeps
    for (uint32_t f = 0; f < lclpep; f++) {
        uint32_t idd = generate_index(f, lclpep, nodes, myid, policy);
        hcx_pep_entry e;
        std::memset(&e, 0, sizeof e);
        e.mass = orc_pep_mass(seqs + (size_t)idd * len, len);
        e.seq_id = idd;
        out[f] = e;
    }
    // mods.cpp:448-496 MODS_GenerateMods
    if (lclmod > 0) {
        hcx_pep_entry *mod_entries = out + lclpep;
        auto keep = [&](uint32_t key) {
            return policy == HCX_POLICY_CHUNK ? (key / lclmod) == myid : key % nodes == myid;
        };
        std::vector<hcx_pep_entry> vars;
        for (uint32_t i = 0; i < nseqs; i++) {
            if (!(var_count[i + 1] - var_count[i])) continue;
            uint32_t global = var_count[i];
            uint32_t local = var_count[i] / nodes + (myid < var_count[i] % nodes ? 1 : 0);
            std::vector<int> cond = m.cond;
            hcx_pep_entry cont;
            std::memset(&cont, 0, sizeof cont);
            vars.clear();
            const char *s = seqs + (size_t)i * len;
            mod_list(s, len, cond, m.limit, cont, 0, false, 0, i, global, vars, m, keep);
            for (size_t k = 0; k < vars.size(); k++) {
                vars[k].mass = orc_mod_mass(s, len, vars[k].mod_num, p);
                mod_entries[local + k] = vars[k];
            }
            std::qsort(mod_entries + local, vars.size(), sizeof(hcx_pep_entry), cmp_var_entries);
        }
    }
    // lbe_internal.cpp:184-187
    if (sort) std::sort(out, out + lclpep + lclmod, cmp_pep_entries);
    return HCX_SUCCESS;
}

extern "C" void orc_std_sort_mass(float *keys, uint32_t *vals, uint32_t n) {
    // std::sort of pepEntry-like records by mass alone: the element type does
    // not enter the algorithm's steps, only the comparator does
    std::vector<hcx_pep_entry> e(n);
    for (uint32_t i = 0; i < n; i++) {
        std::memset(&e[i], 0, sizeof e[i]);
        e[i].mass = keys[i];
        e[i].seq_id = vals[i];
    }
    std::sort(e.begin(), e.end(), cmp_pep_entries);
    for (uint32_t i = 0; i < n; i++) {
        keys[i] = e[i].mass;
        vals[i] = e[i].seq_id;
    }
}

extern "C" void orc_qsort_sites(uint64_t *sites, uint32_t *vals, uint32_t n) {
    std::vector<hcx_pep_entry> e(n);
    for (uint32_t i = 0; i < n; i++) {
        std::memset(&e[i], 0, sizeof e[i]);
        e[i].sites = sites[i];
        e[i].seq_id = vals[i];
    }
    std::qsort(e.data(), n, sizeof(hcx_pep_entry), cmp_var_entries);
    for (uint32_t i = 0; i < n; i++) {
        sites[i] = e[i].sites;
        vals[i] = e[i].seq_id;
    }
}

// M. D. McIlroy, "A killer adversary for quicksort", Software: Practice and
// Experience 29(4), 1999: std::sort runs on item ids with a comparator that
// leaves items "gas" (larger than every solid value) until one of two gas
// items must be decided, freezing the one that is not the pivot candidate.
// The frozen values are an input on which the same sort makes the same
// comparisons, driving its partitions to their worst case (and std::sort to
// its heap-sort fallback).  Items still gas at the end keep the value n - 1.
extern "C" void orc_antiqsort(uint32_t n, float *keys) {
    std::vector<uint32_t> val(n, n - 1), ids(n);
    const uint32_t gas = n - 1;
    uint32_t nsolid = 0, candidate = 0;
    for (uint32_t i = 0; i < n; i++) ids[i] = i;
    auto less = [&](uint32_t x, uint32_t y) {
        if (val[x] == gas && val[y] == gas) {
            if (x == candidate) val[x] = nsolid++;
            else val[y] = nsolid++;
        }
        if (val[x] == gas) candidate = x;
        else if (val[y] == gas) candidate = y;
        return val[x] < val[y];
    };
    std::sort(ids.begin(), ids.end(), less);
    for (uint32_t i = 0; i < n; i++) keys[i] = (float)val[i];
}
