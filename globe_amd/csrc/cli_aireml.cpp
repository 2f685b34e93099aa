// runAIREML: drop-in for the reference CLI (varCompEst.cpp:1347-1500) on one MI355X.
//
//   runAIREML <input_filename> <grm_prefix> <output_prefix> [--maxiter n] [--tol v]
//             [--numMC n] [--sig2e v] [--sig2g v] [--sig2b0 v] [--sigb01 v] [--sig2b1 v]
// reads <grm_prefix>.grm.bin, tees everything it prints to <output_prefix>.log and
// writes no other file, as the reference. Unknown flags are ignored (as :113-132).
// Opt-in extensions: --probes gaussian|rademacher|exact (default gaussian: the
// reference's tilde-V probes), --seed n, --pcgTol v (1e-5), --precision
// fp64|mixed|fp64-valu, --device d, --geno <plinkPrefix> (packed genotypes instead of
// the GRM; the reference's runAIREML only reads a GRM), --writeSolutions (also write
// <output_prefix>_beta.bin, _uhat.bin and _bhat.bin: the GLS / BLUP solve at the fitted
// components, the files the reference's commented-out saveVector calls name,
// varCompEst.cpp:1493-1495), --loglik n|exact (also print the REML log-likelihood at the
// fitted components: log|V| by Lanczos quadrature over n Rademacher probe solves, or
// exactly over every column of the preconditioner's factor; the reference prints none).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/globe.h"
#include "io.h"

namespace {

std::ofstream g_log;

void emit(const std::string& s) {
    std::cout << s << std::endl;
    if (g_log.is_open()) g_log << s << std::endl;
}

void log_cb(const char* line, void*) { emit(line); }

[[noreturn]] void die(const std::string& msg) {
    std::cerr << msg << std::endl;
    if (g_log.is_open()) g_log << msg << std::endl;
    std::exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "Usage: " << argv[0]
                  << " <input_filename> <grm_prefix> <output_prefix> [--maxiter: optional] [--tol: optional]"
                     " [--numMC: optional] [--sig2e: optional] [--sig2g: optional] [--sig2b0: optional]"
                     " [--sigb01: optional] [--sig2b1: optional]"
                  << std::endl;
        return EXIT_FAILURE;
    }
    const std::string input = argv[1], grm_prefix = argv[2], out = argv[3];
    globe_aireml_opts o{};
    o.max_iter = 50;
    o.tol = 1e-3;
    o.num_mc = 15;
    o.probes = 0;
    o.seed = 20261203ull;
    o.pcg = globe_pcg_opts{1e-5, 1000, GLOBE_X0_PRECOND, GLOBE_RULE_ALL_COLUMNS, 0, GLOBE_PREC_FP64, 0};
    double theta[5] = {0.5, 0.3, 0.2, 0.1, 0.1};  // VarCompDataStruct defaults (:47-51)
    int device = 0;
    bool write_solutions = false;
    int loglik_probes = 0;  // 0: none, -1: exact
    std::string geno;
    for (int i = 4; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) die("[Error] Flag " + a + " expects an argument.");
            return argv[++i];
        };
        if (a == "--maxiter") o.max_iter = std::stoi(next());
        else if (a == "--tol") o.tol = std::stof(next());
        else if (a == "--numMC") o.num_mc = std::stoi(next());
        else if (a == "--sig2e") theta[0] = std::stof(next());
        else if (a == "--sig2g") theta[1] = std::stof(next());
        else if (a == "--sig2b0") theta[2] = std::stof(next());
        else if (a == "--sigb01") theta[3] = std::stof(next());
        else if (a == "--sig2b1") theta[4] = std::stof(next());
        else if (a == "--seed") o.seed = std::stoull(next());
        else if (a == "--pcgTol") o.pcg.tol = std::stod(next());
        else if (a == "--device") device = std::stoi(next());
        else if (a == "--geno") geno = next();
        else if (a == "--writeSolutions") write_solutions = true;
        else if (a == "--loglik") {
            const std::string v = next();
            loglik_probes = v == "exact" ? -1 : std::atoi(v.c_str());
            if (!(loglik_probes == -1 || loglik_probes >= 2))
                die("[Error] --loglik takes a probe count >= 2 or exact");
        }
        else if (a == "--probes") {
            const std::string v = next();
            if (v == "gaussian") o.probes = 0;
            else if (v == "rademacher") o.probes = 1;
            else if (v == "exact") o.probes = 2;
            else die("[Error] --probes must be gaussian, rademacher or exact");
        } else if (a == "--precision") {
            const std::string v = next();
            if (v == "fp64") o.pcg.precision = GLOBE_PREC_FP64;
            else if (v == "mixed") o.pcg.precision = GLOBE_PREC_MIXED;
            else if (v == "fp64-valu") o.pcg.precision = GLOBE_PREC_FP64_VALU;
            else die("[Error] --precision must be fp64, mixed or fp64-valu");
        }
    }
    g_log.open(out + ".log");
    std::string cmd = "Command: \n";
    for (int i = 0; i < argc; ++i) cmd += std::string(argv[i]) + " ";
    emit(cmd + "\n");

    globe_io::Cohort c;
    try {
        c = globe_io::read_input(input);
    } catch (const std::exception& e) {
        die(e.what());
    }
    emit("Found N = " + std::to_string(c.n) + " indivuduals, p = " + std::to_string(c.p) +
         " fixed effects (in addition to TIME), and Nrec = " + std::to_string(c.nrec) +
         " records in the input file.");
    globe_ctx* ctx = nullptr;
    if (globe_create(device, &ctx) != GLOBE_OK) die("[Error] no usable GPU");
    auto check = [&](int rc, const char* what) {
        if (rc != GLOBE_OK) die(std::string("[Error] ") + what + ": " + globe_last_error(ctx));
    };
    check(globe_set_cohort(ctx, c.nrec, c.n, c.p, c.indiv.data(), c.times.data(), c.X.data(), c.nrec),
          "cohort");
    if (geno.empty()) {
        std::vector<float> lt;
        try {
            lt = globe_io::read_grm(grm_prefix, c.n);
        } catch (const std::exception& e) {
            die(e.what());
        }
        check(globe_load_grm(ctx, lt.data()), "GRM");
        emit("GRM read successfully.");
    } else {
        globe_io::BedInfo bed;
        try {
            bed = globe_io::check_bed(geno, c.n);
        } catch (const std::exception& e) {
            die(e.what());
        }
        int64_t kept = 0;
        check(globe_load_bed_file(ctx, geno.c_str(), 0.01f, 0.2f, nullptr, &kept), "genotypes");
        emit("Genotypes staged: " + std::to_string(kept) + " / " + std::to_string(bed.m) + " SNPs kept.");
    }
    emit("Input files read and processed successfully.");
    double se[5], ai[25], spi = 0.0;
    int32_t iters = 0;
    check(globe_aireml_fit(ctx, c.y.data(), &o, theta, se, ai, &iters, &spi, log_cb, nullptr), "AIREML");
    char buf[256];
    std::snprintf(buf, sizeof(buf), "AIREML: %d iterations, %.4f s per iteration", iters, spi);
    emit(buf);
    if (write_solutions) {
        // fitMME's solve at the fitted components (x0 = Q^-1 b, all-columns rule)
        check(globe_set_varcomp(ctx, theta), "variance components");
        std::vector<double> beta(c.p + 1), uhat(c.n), bhat(2 * (size_t)c.n);
        globe_pcg_stats st{};
        check(globe_gls_blup(ctx, c.y.data(), &o.pcg, beta.data(), nullptr, uhat.data(), bhat.data(), &st),
              "solutions");
        globe_io::save_vector(beta, out + "_beta.bin");
        globe_io::save_vector(uhat, out + "_uhat.bin");
        globe_io::save_vector(bhat, out + "_bhat.bin");
        emit("Solutions written to " + out + "_beta.bin, _uhat.bin and _bhat.bin.");
    }
    if (loglik_probes != 0) {
        check(globe_set_varcomp(ctx, theta), "variance components");
        globe_pcg_opts lo = o.pcg;
        lo.max_iter = 1000;
        globe_loglik ll{};
        check(globe_reml_loglik(ctx, c.y.data(), loglik_probes < 0 ? 2 : 1, loglik_probes, o.seed, &lo, &ll),
              "log-likelihood");
        std::snprintf(buf, sizeof(buf),
                      "REML log-likelihood at the estimates: %.10g (log|V| = %.10g +/- %.3g, %d probes, %.3f s)",
                      ll.loglik, ll.logdet_v, ll.logdet_v_se, ll.probes, ll.seconds);
        emit(buf);
    }
    globe_destroy(ctx);
    g_log.close();
    return 0;
}
