// makeGRM: writes the GRM G = W W^T / M of a PLINK cohort as a GCTA <output>.grm.bin
// (n(n+1)/2 float32, lower triangle row by row), the file runPCGsolver --grm and
// runAIREML read. The reference declares a GRM builder (computeFullGRM,
// PCGsolver.cpp:165-167) but ships none; W is its standardised genotype matrix
// (getStdGenotype, :831-842) over the SNPs its QC keeps (prepGenotypeInput, :615-763).
//
//   makeGRM --geno <plinkPrefix> --output <prefix> [--minMAF v] [--maxMissingRate v]
//           [--device d] [--debug]
// n is the line count of <plinkPrefix>.fam (the individuals in .bed order). The GRM is
// built on the device (exact int8-digit products, rounded to FP32) and streamed to the
// file in chunks of rows, so it never has to fit in host memory at once.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "../../include/globe.h"
#include "io.h"

namespace {

// the partial output file, removed when the run fails
std::string g_partial;

[[noreturn]] void die(const std::string& msg) {
    std::cerr << msg << std::endl;
    if (!g_partial.empty()) std::remove(g_partial.c_str());
    std::exit(EXIT_FAILURE);
}

void check(globe_ctx* ctx, int rc, const char* what) {
    if (rc != GLOBE_OK) die(std::string("[Error] ") + what + ": " + globe_last_error(ctx));
}

}  // namespace

int main(int argc, char** argv) {
    std::string geno, output;
    float min_maf = 0.01f, max_miss = 0.05f;
    int device = 0;
    bool debug = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) die("[Error] Flag " + a + " expects an argument.");
            return argv[++i];
        };
        if (a == "--geno") geno = next();
        else if (a == "--output") output = next();
        else if (a == "--minMAF") min_maf = std::stof(next());
        else if (a == "--maxMissingRate") max_miss = std::stof(next());
        else if (a == "--device") device = std::stoi(next());
        else if (a == "--debug") debug = true;
        else die("[Error] Unknown flag: " + a);
    }
    if (geno.empty() || output.empty())
        die("Usage: " + std::string(argv[0]) +
            " --geno <plinkPrefix> --output <prefix> [--minMAF v] [--maxMissingRate v] [--device d]");
    if (min_maf < 0.f || min_maf > 0.5f) die("[Error] Invalid MAF threshold: must be in [0, 0.5].");
    if (max_miss < 0.f || max_miss > 1.f) die("[Error] Invalid missing rate threshold: must be in [0, 1].");

    // every input is checked before the device is touched
    int64_t n64 = 0;
    {
        std::ifstream fm(geno + ".fam");
        if (!fm.is_open()) die("[Error] Error opening .fam file: " + geno + ".fam");
        n64 = std::count(std::istreambuf_iterator<char>(fm), std::istreambuf_iterator<char>(), '\n');
    }
    if (n64 <= 0 || n64 > (1ll << 31) - 1) die("[Error] " + geno + ".fam lists no samples");
    const int32_t n = (int32_t)n64;
    globe_io::BedInfo bed;
    try {
        bed = globe_io::check_bed(geno, n);
    } catch (const std::exception& e) {
        die(e.what());
    }
    // written under a temporary name and renamed once complete, so a failed run never
    // leaves a truncated GRM under the final name
    const std::string path = output + ".grm.bin";
    const std::string tmp = path + ".tmp";
    std::ofstream f(tmp, std::ios::binary);
    if (!f.is_open()) die("[Error] Error opening output file: " + tmp);
    g_partial = tmp;

    globe_ctx* ctx = nullptr;
    if (globe_create(device, &ctx) != GLOBE_OK)
        die("[Error] no usable GPU (device " + std::to_string(device) + ", " +
            std::to_string(globe_device_count()) + " visible)");
    // a cohort of one record per individual: the builder only needs n
    std::vector<int32_t> indiv(n);
    for (int32_t i = 0; i < n; ++i) indiv[i] = i;
    std::vector<double> zeros((size_t)n, 0.0);
    check(ctx, globe_set_cohort(ctx, n, n, 0, indiv.data(), zeros.data(), zeros.data(), n), "cohort");
    int64_t kept = 0;
    const auto t0 = std::chrono::steady_clock::now();
    check(ctx, globe_load_bed_file(ctx, geno.c_str(), min_maf, max_miss, nullptr, &kept), "genotypes");
    std::cout << "[makeGRM] " << n << " individuals, " << kept << " / " << bed.m << " SNPs kept." << std::endl;
    if (kept == 0) die("[Error] no SNP passes QC");
    check(ctx, globe_grm_from_geno(ctx), "GRM");
    const double t_build = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // stream the triangle out in chunks of about 64M floats
    std::vector<float> buf;
    for (int64_t r0 = 0; r0 < n;) {
        int64_t r1 = r0 + 1;
        while (r1 < n && (r1 + 1) * (r1 + 2) / 2 - r0 * (r0 + 1) / 2 <= (64ll << 20)) ++r1;
        const size_t cnt = (size_t)(r1 * (r1 + 1) / 2 - r0 * (r0 + 1) / 2);
        buf.resize(cnt);
        check(ctx, globe_get_grm_rows(ctx, r0, r1, buf.data()), "GRM read-out");
        f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(cnt * sizeof(float)));
        if (!f) die("[Error] Error writing " + tmp);
        r0 = r1;
    }
    f.close();
    if (!f) die("[Error] Error writing " + tmp);
    if (std::rename(tmp.c_str(), path.c_str()) != 0) die("[Error] Error renaming " + tmp + " to " + path);
    g_partial.clear();
    const double t_all = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "[makeGRM] Wrote " << path << " (" << (size_t)n * (n + 1) / 2 * sizeof(float) << " bytes)";
    if (debug) std::cout << ": ingest + build " << t_build << " s, total " << t_all << " s";
    std::cout << std::endl;
    globe_destroy(ctx);
    return 0;
}
