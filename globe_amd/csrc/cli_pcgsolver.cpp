// runPCGsolver: drop-in for the reference CLI (PCGsolver.cpp:235-306) on one MI355X.
//
// Same flags, defaults, input files and output files:
//   --input <file> --output <prefix> --sig2e v --sig2g v --sig2b0 v --sigb01 v --sig2b1 v
//   (--geno <plinkPrefix> | --grm <grmPrefix>) [--debug] [--minMAF v] [--maxMissingRate v]
// writes <prefix>_beta.bin (p+1, TIME last), <prefix>_uhat.bin (N), <prefix>_bhat.bin
// (2N, [b0, b1] per individual), each int32 n + n float32.
// Opt-in extensions (absent from the reference): --pcgTol (default 1e-5, the
// reference's hard-coded value), --maxIter (1000), --precision fp64|mixed|fp64-valu,
// --device.
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/globe.h"
#include "io.h"

namespace {

[[noreturn]] void die(const std::string& msg) {
    std::cerr << msg << std::endl;
    std::exit(EXIT_FAILURE);
}

void usage(const char* argv0) {
    std::cerr << "Usage:\n"
              << argv0 << " --input <file> --output <prefix> "
              << " --sig2e <val> --sig2g <val> --sig2b0 <val> --sigb01 <val> --sig2b1 <val>\n"
              << "  EITHER:  --geno <plinkPrefix>\n"
              << "      OR:  --grm <grmPrefix>\n"
              << " Optional:\n"
              << "   [--debug] [--minMAF val] [--maxMissingRate val]\n"
              << "   [--pcgTol val] [--maxIter n] [--precision fp64|mixed|fp64-valu] [--device d]\n";
    std::exit(EXIT_FAILURE);
}

void check(globe_ctx* c, int rc, const char* what) {
    if (rc != GLOBE_OK) die(std::string("[Error] ") + what + ": " + globe_last_error(c));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) usage(argv[0]);
    bool debug = false, have_in = false, have_out = false, hs[5] = {false, false, false, false, false};
    float min_maf = 0.01f, max_miss = 0.2f;  // GlobalOptions defaults (PCGsolver.cpp:32-37)
    float sig[5] = {0.1f, 0.1f, 0.1f, 0.1f, 0.1f};
    std::string input, output, geno, grm;
    double tol = 1e-5;
    int max_iter = 1000, device = 0, precision = GLOBE_PREC_FP64;
    const char* names[5] = {"--sig2e", "--sig2g", "--sig2b0", "--sigb01", "--sig2b1"};
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) die("[Error] Flag " + arg + " expects an argument.");
            return std::string(argv[++i]);
        };
        bool matched = false;
        for (int q = 0; q < 5; ++q)
            if (arg == names[q]) {
                sig[q] = std::stof(next());
                hs[q] = true;
                matched = true;
            }
        if (matched) continue;
        if (arg == "--debug") debug = true;
        else if (arg == "--minMAF") min_maf = std::stof(next());
        else if (arg == "--maxMissingRate") max_miss = std::stof(next());
        else if (arg == "--geno") geno = next();
        else if (arg == "--grm") grm = next();
        else if (arg == "--input") { input = next(); have_in = true; }
        else if (arg == "--output") { output = next(); have_out = true; }
        else if (arg == "--pcgTol") tol = std::stod(next());
        else if (arg == "--maxIter") max_iter = std::stoi(next());
        else if (arg == "--device") device = std::stoi(next());
        else if (arg == "--precision") {
            const std::string v = next();
            if (v == "fp64") precision = GLOBE_PREC_FP64;
            else if (v == "mixed") precision = GLOBE_PREC_MIXED;
            else if (v == "fp64-valu") precision = GLOBE_PREC_FP64_VALU;
            else die("[Error] --precision must be fp64, mixed or fp64-valu");
        } else die("[Error] Unknown flag: " + arg);
    }
    if (!have_in || !have_out || !hs[0] || !hs[1] || !hs[2] || !hs[3] || !hs[4])
        die("[Error] Missing required flags.");
    if (min_maf < 0 || min_maf > 0.5f || max_miss < 0 || max_miss > 1.0f)
        die("[Error] Invalid MAF or missing rate thresholds.");
    if (geno.empty() == grm.empty())
        die("[Error] You must specify exactly one of:\n  --geno <plinkPrefix>\nOR\n  --grm  <grmPrefix>\nbut not both.");

    globe_io::Cohort c;
    try {
        c = globe_io::read_input(input);
    } catch (const std::exception& e) {
        die(e.what());
    }
    if (debug)
        std::cout << "[checkInputs] Found:\n   N = " << c.n << " unique individuals\n   p = " << c.p
                  << " fixed effects (not counting TIME)\n   Nrec = " << c.nrec << " records total." << std::endl;

    // open every input before touching the device, so file errors exit like the reference
    globe_io::BedInfo bed;
    std::vector<float> lt;
    try {
        if (!geno.empty()) bed = globe_io::check_bed(geno, c.n);
        else lt = globe_io::read_grm(grm, c.n);
    } catch (const std::exception& e) {
        die(e.what());
    }
    globe_ctx* ctx = nullptr;
    if (globe_create(device, &ctx) != GLOBE_OK)
        die("[Error] no usable GPU (device " + std::to_string(device) + ", " +
            std::to_string(globe_device_count()) + " visible)");
    check(ctx, globe_set_cohort(ctx, c.nrec, c.n, c.p, c.indiv.data(), c.times.data(), c.X.data(), c.nrec),
          "cohort");
    if (!geno.empty()) {
        int64_t kept = 0;
        check(ctx, globe_load_bed_file(ctx, geno.c_str(), min_maf, max_miss, nullptr, &kept), "genotypes");
        std::cout << "[prepGenotypeInput] Processed " << bed.m << " / " << bed.m << " SNPs." << std::endl;
        if (debug) std::cout << "[prepGenotypeInput] Filtered SNPs = " << kept << " / " << bed.m << std::endl;
    } else {
        check(ctx, globe_load_grm(ctx, lt.data()), "GRM");
    }
    std::cout << "[prepareAllInputs] Input prepared successfully." << std::endl;
    const double theta[5] = {sig[0], sig[1], sig[2], sig[3], sig[4]};
    check(ctx, globe_set_varcomp(ctx, theta), "variance components");

    globe_pcg_opts o{tol, max_iter, GLOBE_X0_ZERO, GLOBE_RULE_MAX_RATIO, 1, precision, 0};
    globe_pcg_stats st{};
    std::vector<double> beta(c.p + 1), uhat(c.n), bhat(2 * (size_t)c.n);
    std::cout << "[runPCG_multi] Starting PCG solver..." << std::endl;
    auto t0 = std::chrono::steady_clock::now();
    check(ctx, globe_gls_blup(ctx, c.y.data(), &o, beta.data(), nullptr, uhat.data(), bhat.data(), &st),
          "solve");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st.stalled)
        std::cout << "[runPCG_multi] Stalled (no significant improvement for 10 iters)." << std::endl;
    std::cout << "[runPCG_multi] Converged at iteration " << st.iters << std::endl;
    if (debug) {
        std::cout << "[runPCG_multi] max |r|/|b| = " << st.max_ratio << ", PCG " << st.seconds
                  << " s on the device, estimators " << secs << " s" << std::endl;
        std::cout << "beta:";
        for (double b : beta) std::cout << " " << b;
        std::cout << std::endl;
    }
    globe_io::save_vector(beta, output + "_beta.bin");
    globe_io::save_vector(uhat, output + "_uhat.bin");
    globe_io::save_vector(bhat, output + "_bhat.bin");
    globe_destroy(ctx);
    std::cout << "[main] Done." << std::endl;
    return 0;
}
