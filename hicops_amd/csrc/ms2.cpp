// ms2.cpp -- MS2 text ingest (include/hcx_ms2.h): the reference's file
// pre-scan (MSQuery::initialize, msquery.cpp:96-170) and one
// MSQuery::readspectrum (msquery.cpp:300-505) per spectrum, over a whole file.
#include "../../include/hcx_ms2.h"
#include "../../include/hcx.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

struct hcx_ms2 {
    std::vector<uint32_t> mz, in, idx;
    std::vector<double> prec, rt;
    std::vector<int32_t> z;
    uint32_t maxlen = 0;
};

namespace {

// The fields of MSQuery::spectrum that carry over from one spectrum to the
// next (msquery.cpp:40-42 start them at 0).
struct SpecState {
    int32_t z = 0;
    double prec = 0.0, rt = 0.0;
};

// strtok_r on a copy of the line, as the reference does on line.c_str()
struct Tokens {
    std::vector<char> buf;
    char *save = nullptr;
    bool first = true;
    Tokens(const std::string &line) : buf(line.begin(), line.end()) { buf.push_back('\0'); }
    const char *next(const char *delim) {
        char *t = strtok_r(first ? buf.data() : nullptr, delim, &save);
        first = false;
        return t;
    }
};

// msquery.cpp:325-345 (and :424-446): "Z <charge> <[M+H]+>"
void z_line(const std::string &line, SpecState &st) {
    Tokens t(line);
    t.next(" \t");
    const char *mh = t.next(" \t");
    std::string val = mh ? mh : "1";
    st.z = std::atoi(val.c_str());
    mh = t.next(" \t");
    val = mh ? mh : "0.01";
    st.prec = (double)std::atof(val.c_str());
}

// msquery.cpp:346-370: "I RTime <t>" (other I lines are ignored)
void i_line(const std::string &line, SpecState &st) {
    Tokens t(line);
    t.next(" \t");
    const char *mh = t.next(" \t");
    std::string val = mh ? mh : "";
    if (val.compare("RTime") == 0) {
        val = "0.00";
        mh = t.next(" \t");
        if (mh) val = mh;
        st.rt = (double)std::atof(val.c_str());
    }
}

// msquery.cpp:385-406 (first spectrum: missing tokens read "0.01") and
// :478-500 (later spectra: "0.0"); peak lines split on spaces only
void peak_line(const std::string &line, bool first_spec, uint32_t scale, hcx_ms2 *m) {
    Tokens t(line);
    const char *mz1 = t.next(" ");
    const char *in1 = t.next(" ");
    std::string mz = first_spec ? "0.01" : "0.0";
    std::string in = first_spec ? "0.01" : "0.0";
    if (mz1) mz = mz1;
    if (in1) in = in1;
    m->mz.push_back((uint32_t)((double)std::atof(mz.c_str()) * scale));
    m->in.push_back((uint32_t)((double)std::atof(in.c_str()) * 1000));
}

}  // namespace

extern "C" int hcx_ms2_read(const char *path, uint32_t scale, hcx_ms2 **out) {
    if (!path || !out) return HCX_ERR_INVLD_PTR;
    // msquery.cpp:96-170: count the S lines and the largest spectrum
    uint64_t count = 0;
    uint32_t largest = 0, size = 0;
    {
        std::ifstream f(path);
        if (!f.is_open()) return HCX_ERR_FILE_NOT_FOUND;
        std::string line;
        while (!f.eof()) {
            std::getline(f, line);
            if (line.empty()) continue;
            if (line[0] == 'S') {
                count++;
                largest = std::max(size, largest);
                size = 0;
            } else if (line[0] == 'H' || line[0] == 'I' || line[0] == 'D' || line[0] == 'Z') {
                continue;
            } else {
                size++;
            }
        }
        largest = std::max(size, largest);
        if (largest < 1) return HCX_ERR_INVLD_SIZE;
        if (count > 0x7FFFFFFFull) return HCX_ERR_INVLD_SIZE;
    }
    hcx_ms2 *m = new (std::nothrow) hcx_ms2;
    if (!m) return HCX_ERR_BAD_MEM_ALLOC;
    m->maxlen = largest;
    std::ifstream f(path);
    if (!f.is_open()) {
        delete m;
        return HCX_ERR_FILE_NOT_FOUND;
    }
    SpecState st;
    std::string line;
    m->idx.push_back(0);
    // one readspectrum call per counted spectrum (MSQuery::extractbatch,
    // msquery.cpp:280-286)
    for (uint64_t spec = 0; spec < count; spec++) {
        const bool first = spec == 0;
        bool scan = false;
        while (!f.eof()) {
            std::getline(f, line);
            if (line.empty() || line[0] == 'H' || line[0] == 'D') continue;
            if (line[0] == 'Z') {
                z_line(line, st);
            } else if (line[0] == 'I') {
                i_line(line, st);
            } else if (line[0] == 'S') {
                // the first spectrum starts at the first S and ends at the
                // second (msquery.cpp:372-381); later ones end at the next S
                if (first && !scan) {
                    scan = true;
                    continue;
                }
                break;
            } else {
                peak_line(line, first, scale, m);
            }
        }
        if (m->mz.size() > 0xFFFFFFFFull) {
            delete m;
            return HCX_ERR_INVLD_SIZE;
        }
        m->idx.push_back((uint32_t)m->mz.size());
        m->prec.push_back(st.prec);
        m->z.push_back(st.z);
        m->rt.push_back(st.rt);
    }
    *out = m;
    return HCX_SUCCESS;
}

extern "C" int hcx_ms2_free(hcx_ms2 *m) {
    delete m;
    return HCX_SUCCESS;
}

extern "C" int hcx_ms2_info(const hcx_ms2 *m, int32_t *nspec, uint64_t *npeaks, uint32_t *maxlen) {
    if (!m) return HCX_ERR_INVLD_PTR;
    if (nspec) *nspec = (int32_t)m->prec.size();
    if (npeaks) *npeaks = m->mz.size();
    if (maxlen) *maxlen = m->maxlen;
    return HCX_SUCCESS;
}

extern "C" int hcx_ms2_copy(const hcx_ms2 *m, uint32_t *mz, uint32_t *intensity, uint32_t *idx, double *prec_mz,
                            int32_t *z, double *rtime) {
    if (!m) return HCX_ERR_INVLD_PTR;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(mz, m->mz.data(), m->mz.size() * 4);
    cp(intensity, m->in.data(), m->in.size() * 4);
    cp(idx, m->idx.data(), m->idx.size() * 4);
    cp(prec_mz, m->prec.data(), m->prec.size() * 8);
    cp(z, m->z.data(), m->z.size() * 4);
    cp(rtime, m->rt.data(), m->rt.size() * 8);
    return HCX_SUCCESS;
}
