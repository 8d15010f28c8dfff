// dslim_driver.cpp -- drives the reference-shaped C++ entry points the way
// the reference application does (source/app/hicops.cpp:280-310 builds one
// Index per length with DSLIM_Construct, then DSLIM_InitializeScorecard;
// DSLIM_SearchManager feeds Queries batches to DSLIM_QuerySpectrum,
// dslim_query.cpp:160-400), and records every PSM that reaches
// DFile_PrintScore.
//
// usage: dslim_driver <dir> <batches>
//   <dir>/params.txt   "key value" lines (gParams fields + limits)
//   <dir>/lengths.txt  one peptide length per line
//   <dir>/seqs_<L>.bin, <dir>/ents_<L>.bin   sequences, pepEntry[24 B]
//   <dir>/moz.bin inten.bin idx.bin prec.bin chg.bin   CSR spectra
// With "lbe 1" in params.txt the database comes from the peptide database
// stage instead: <dir>/fasta_<L>.txt (one peptide per line, '>' headers) and
// <dir>/mods.txt (params.modconditions) through MODS_Initialize,
// LBE_CountPeps, LBE_CreatePartitions, LBE_Initialize and LBE_Distribute, as
// hicops.cpp:200-260 calls them; <dir>/lbe_ents_<L>.bin and lbe_seqs_<L>.bin
// receive the resulting pepEntries and sequences.
// writes <dir>/psms.txt: specid psid idxoffset sharedions totalions
//                        hyperscore-bits e_x npsms pchg fileIndex
// Timing: "repeat R" in params.txt runs the batches R times, "quiet 1" makes
// the sink count PSMs instead of printing them; <dir>/timing.txt receives the
// wall time of every DSLIM_QuerySpectrum call (ms, one per line) and the
// PSMs reported per pass.
// Built with -DHCX_TSV_DRIVER (tests/cpp/tsv_driver) it instead leaves the
// PSMs to the library's TSV writer: DFile_InitFiles with params.workspace =
// <dir> and two query threads, the batches, DFile_DeinitFiles; <dir>/rt.bin
// holds retention times.
#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/hcx_dslim.hpp"

gParams params;   // the application's parameter block, as in source/app/hicops.cpp:31
static uint64_t g_psms = 0;   // PSMs reported through this program's sink

#ifdef HCX_TSV_DRIVER
// the application's query file list (source/app/hicops.cpp:28)
std::vector<std::string> queryfiles = {"q0.ms2", "q1.ms2", "q2.ms2", "sample_3.ms2"};
#else
static FILE *g_out = nullptr;
static bool g_quiet = false;

// The application's result sink (replaces the library's own definition).
status_t DFile_PrintScore(Index *, uint_t specid, float pmass, hCell *psm, double e_x, uint_t npsms) {
    g_psms++;
    if (g_quiet) return SLM_SUCCESS;
    uint32_t hb;
    std::memcpy(&hb, &psm->hyperscore, 4);
    (void)pmass;
    std::fprintf(g_out, "%u %d %u %u %u %08x %.17g %u %d %u\n", specid, psm->psid, psm->idxoffset, psm->sharedions,
                 psm->totalions, hb, e_x, npsms, psm->pchg, psm->fileIndex);
    return SLM_SUCCESS;
}
#endif

template <typename T>
static std::vector<T> load(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        std::fprintf(stderr, "missing %s\n", path.c_str());
        std::exit(2);
    }
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<T> v((size_t)n / sizeof(T));
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return v;
}

static void check(status_t rc, const char *what) {
    if (rc != SLM_SUCCESS) {
        std::fprintf(stderr, "%s failed: %d (%s)\n", what, rc, hcx_last_error());
        std::exit(1);
    }
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <dir> <batches>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int batches = std::atoi(argv[2]);
    std::map<std::string, double> kv;
    {
        std::ifstream f(dir + "/params.txt");
        std::string k;
        double v;
        while (f >> k >> v) kv[k] = v;
    }
    params.maxz = (uint_t)kv["maxz"];
    params.scale = (uint_t)kv["scale"];
    params.min_mass = (uint_t)kv["min_mass"];
    params.max_mass = (uint_t)kv["max_mass"];
    params.dF = (uint_t)kv["dF"];
    params.min_shp = (uint_t)kv["min_shp"];
    params.min_cpsm = (uint_t)kv["min_cpsm"];
    params.topmatches = (uint_t)kv["topmatches"];
    params.dM = kv["dM"];
    params.expect_max = kv["expect_max"];
    params.vModInfo.num_vars = (ushort_t)kv["nmods"];
    for (int i = 0; i < params.vModInfo.num_vars; i++)
        params.vModInfo.vmods[i].modMass = (uint_t)kv["mod" + std::to_string(i)];
    check(DSLIM_SetLimits((uint_t)kv["max_hyperscore"], (uint_t)kv["max_shdpeaks"]), "DSLIM_SetLimits");
    if (kv.count("kernel")) check(DSLIM_SetKernel((int_t)kv["kernel"]), "DSLIM_SetKernel");

    std::vector<int> lengths;
    {
        std::ifstream f(dir + "/lengths.txt");
        int L;
        while (f >> L) lengths.push_back(L);
    }
    std::vector<Index> index(lengths.size());
    std::vector<std::vector<char>> seqs(lengths.size());
    std::vector<std::vector<pepEntry>> ents(lengths.size());
    const bool lbe = kv.count("lbe") && kv["lbe"] != 0;
    if (lbe) {
        std::ifstream mf(dir + "/mods.txt");
        std::getline(mf, params.modconditions);
        params.nodes = 1;
        params.myid = 0;
        check(MODS_Initialize(), "MODS_Initialize");
        for (size_t i = 0; i < lengths.size(); i++) {
            const std::string L = std::to_string(lengths[i]);
            Index &ix = index[i];
            std::memset(&ix, 0, sizeof ix);
            std::string fasta = dir + "/fasta_" + L + ".txt";
            check(LBE_CountPeps(&fasta[0], &ix, (uint_t)lengths[i]), "LBE_CountPeps");
            check(LBE_CreatePartitions(&ix), "LBE_CreatePartitions");
            check(LBE_Initialize(&ix), "LBE_Initialize");
            check(LBE_Distribute(&ix), "LBE_Distribute");
            FILE *f = std::fopen((dir + "/lbe_ents_" + L + ".bin").c_str(), "wb");
            std::fwrite(ix.pepEntries, sizeof(pepEntry), ix.lcltotCnt, f);
            std::fclose(f);
            f = std::fopen((dir + "/lbe_seqs_" + L + ".bin").c_str(), "wb");
            std::fwrite(ix.pepIndex.seqs, 1, ix.pepIndex.AAs, f);
            std::fclose(f);
            ents[i].assign(ix.pepEntries, ix.pepEntries + ix.lcltotCnt);
            check(DSLIM_Construct(&ix), "DSLIM_Construct");
        }
    }
    for (size_t i = 0; i < lengths.size() && !lbe; i++) {
        const std::string L = std::to_string(lengths[i]);
        seqs[i] = load<char>(dir + "/seqs_" + L + ".bin");
        ents[i] = load<pepEntry>(dir + "/ents_" + L + ".bin");
        Index &ix = index[i];
        std::memset(&ix, 0, sizeof ix);
        ix.pepIndex.seqs = seqs[i].data();
        ix.pepIndex.peplen = (ushort_t)lengths[i];
        ix.pepIndex.AAs = (uint_t)seqs[i].size();
        ix.pepEntries = ents[i].data();
        ix.lcltotCnt = ix.totalCount = (uint_t)ents[i].size();
        ix.nChunks = 1;
        ix.chunksize = ix.lastchunksize = ix.lcltotCnt;
        check(DSLIM_Construct(&ix), "DSLIM_Construct");
    }
    check(DSLIM_InitializeScorecard(index.data(), (uint_t)index.size()), "DSLIM_InitializeScorecard");

    auto moz = load<uint32_t>(dir + "/moz.bin");
    auto inten = load<uint32_t>(dir + "/inten.bin");
    auto idx = load<uint32_t>(dir + "/idx.bin");
    auto prec = load<float>(dir + "/prec.bin");
    auto chg = load<int32_t>(dir + "/chg.bin");
    const int nspec = (int)prec.size();
#ifdef HCX_TSV_DRIVER
    auto rt = load<float>(dir + "/rt.bin");
    params.workspace = dir;
    params.threads = 2;
    check(DFile_InitFiles(), "DFile_InitFiles");
#else
    g_out = std::fopen((dir + "/psms.txt").c_str(), "w");
    if (!g_out) return 2;
    g_quiet = kv.count("quiet") && kv["quiet"] != 0;
#endif
    const int repeat = kv.count("repeat") ? std::max(1, (int)kv["repeat"]) : 1;
    std::vector<double> call_ms;
    // the batches of one query file, as the I/O threads hand them over
    for (int rb = 0; rb < repeat * batches; rb++) {
        const int b = rb % batches;
        const int s0 = (int)((long)nspec * b / batches), s1 = (int)((long)nspec * (b + 1) / batches);
        std::vector<uint32_t> bidx(s1 - s0 + 1);
        for (int s = s0; s <= s1; s++) bidx[s - s0] = idx[s] - idx[s0];
        Queries q;
        std::memset(&q, 0, sizeof q);
        q.moz = moz.data() + idx[s0];
        q.intensity = inten.data() + idx[s0];
        q.idx = bidx.data();
        q.precurse = prec.data() + s0;
        q.charges = chg.data() + s0;
#ifdef HCX_TSV_DRIVER
        q.rtimes = rt.data() + s0;
#else
        q.rtimes = nullptr;
#endif
        q.numSpecs = s1 - s0;
        q.numPeaks = (int_t)(idx[s1] - idx[s0]);
        q.batchNum = b;
        q.fileNum = 3;
        const auto t0 = std::chrono::steady_clock::now();
        check(DSLIM_QuerySpectrum(&q, index.data(), (uint_t)index.size()), "DSLIM_QuerySpectrum");
        call_ms.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    {
        FILE *f = std::fopen((dir + "/timing.txt").c_str(), "w");
        for (double m : call_ms) std::fprintf(f, "%.4f\n", m);
        std::fprintf(f, "psms_per_pass %" PRIu64 "\n", g_psms / (uint64_t)repeat);
        std::fclose(f);
    }
#ifdef HCX_TSV_DRIVER
    check(DFile_DeinitFiles(), "DFile_DeinitFiles");
    for (auto &ix : index) check(DSLIM_DeallocateIonIndex(&ix), "DSLIM_DeallocateIonIndex");
    check(DSLIM_DeallocateSC(), "DSLIM_DeallocateSC");
    return 0;
#else
    std::fclose(g_out);
#endif

    // the host copy of one length's index, released with the device index
    check(DSLIM_ExportIonIndex(&index[0]), "DSLIM_ExportIonIndex");
    {
        FILE *f = std::fopen((dir + "/ion0.bin").c_str(), "wb");
        const uint64_t n = (uint64_t)ents[0].size() * 2 * params.maxz * (lengths[0] - 1);
        std::fwrite(index[0].ionIndex[0].iA, 4, n, f);
        std::fwrite(index[0].ionIndex[0].bA, 4, (size_t)params.max_mass * params.scale + 1, f);
        std::fclose(f);
    }
    for (auto &ix : index) check(lbe ? LBE_Deinitialize(&ix) : DSLIM_DeallocateIonIndex(&ix), "deallocate");

    // expeRT::ModelTailFit on a histogram given as survival counts
    {
        const int hb = 10 * (int)kv["max_hyperscore"] + 2;
        std::vector<double> surv(hb, 0.0);
        const double h[] = {5, 10, 8, 6, 3, 1, 0, 1, 1};
        for (int i = 0; i < 9; i++) surv[2 + i] = h[i];
        Results r;
        std::memset(&r, 0, sizeof r);
        r.survival = surv.data();
        r.cpsms = 35;
        r.maxhypscore = 10;
        expeRT e;
        check(e.ModelTailFit(&r), "expeRT::ModelTailFit");
        FILE *f = std::fopen((dir + "/tailfit.txt").c_str(), "w");
        std::fprintf(f, "%d %d %d %d\n", r.mu, r.beta, r.minhypscore, r.nexthypscore);
        std::fclose(f);
    }
    check(DSLIM_DeallocateSC(), "DSLIM_DeallocateSC");
    return 0;
}
