// runAIREML_reducedModel: drop-in for the reference's reduced-model CLI
// (varCompEst_reducedModel_noTime.cpp:967-1099) on one MI355X.
//
//   runAIREML_reducedModel <input_filename> <grm_prefix> <output_prefix> [--maxiter n]
//                          [--tol v] [--numMC n] [--sig2e v] [--sig2g v]
// fits V = sig2g A + sig2e I (one record per individual, no time term) by AI-REML,
// reads <grm_prefix>.grm.bin, tees everything it prints to <output_prefix>.log and
// writes no other file, as the reference. Unknown flags are ignored (as :113-128).
// Opt-in extensions: --probes gaussian|rademacher|exact (default gaussian: the
// reference's tilde-V probes), --seed n, --pcgTol v (1e-5), --precision
// fp64|mixed|fp64-valu, --device d, --loglik n|exact (also print the REML log-likelihood
// at the fitted components: log|V| by Lanczos quadrature over n Rademacher probe solves,
// or exactly; the reference prints none).
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
                  << " <input_filename> <grm_prefix> <output_prefix> [--maxiter <int>] [--tol <float>]"
                     " [--numMC <int>] [--sig2e <float>] [--sig2g <float>]"
                  << std::endl;
        return EXIT_FAILURE;
    }
    const std::string input = argv[1], grm_prefix = argv[2], out = argv[3];
    globe_aireml_opts o{};
    o.max_iter = 50;  // MAX_ITER (:21)
    o.tol = 1e-3;     // TOL (:22)
    o.num_mc = 15;    // NO_MC_SAMPLES (:23)
    o.probes = 0;
    o.seed = 20261203ull;
    o.pcg = globe_pcg_opts{1e-5, 1000, GLOBE_X0_PRECOND, GLOBE_RULE_ALL_COLUMNS, 0, GLOBE_PREC_FP64, 0};
    double theta[2] = {0.5, 0.5};  // VarCompDataStruct defaults (:43-44)
    int device = 0;
    int loglik_probes = 0;  // 0: none, -1: exact
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
        else if (a == "--seed") o.seed = std::stoull(next());
        else if (a == "--pcgTol") o.pcg.tol = std::stod(next());
        else if (a == "--device") device = std::stoi(next());
        else if (a == "--loglik") {
            const std::string v = next();
            loglik_probes = v == "exact" ? -1 : std::atoi(v.c_str());
            if (!(loglik_probes == -1 || loglik_probes >= 2))
                die("[Error] --loglik takes a probe count >= 2 or exact");
        } else if (a == "--probes") {
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
        c = globe_io::read_input_reduced(input);
    } catch (const std::exception& e) {
        die(e.what());
    }
    emit("Found N = " + std::to_string(c.n) + " indivuduals, p = " + std::to_string(c.p + 1) +
         " fixed effects, and Nrec = " + std::to_string(c.nrec) + " records in the input file.");
    std::vector<float> lt;
    try {
        lt = globe_io::read_grm(grm_prefix, c.n);
    } catch (const std::exception& e) {
        die(e.what());
    }
    globe_ctx* ctx = nullptr;
    if (globe_create(device, &ctx) != GLOBE_OK) die("[Error] no usable GPU");
    auto check = [&](int rc, const char* what) {
        if (rc != GLOBE_OK) die(std::string("[Error] ") + what + ": " + globe_last_error(ctx));
    };
    check(globe_set_cohort(ctx, c.nrec, c.n, c.p, c.indiv.data(), c.times.data(), c.X.data(), c.nrec),
          "cohort");
    check(globe_load_grm(ctx, lt.data()), "GRM");
    emit("Input files read and processed successfully.");
    double se[2], ai[4], spi = 0.0;
    int32_t iters = 0;
    check(globe_aireml_fit_reduced(ctx, c.y.data(), &o, theta, se, ai, &iters, &spi, log_cb, nullptr),
          "AIREML");
    char buf[256];
    std::snprintf(buf, sizeof(buf), "AIREML: %d iterations, %.4f s per iteration", iters, spi);
    emit(buf);
    if (loglik_probes != 0) {
        const double th5[5] = {theta[0], theta[1], 0.0, 0.0, 0.0};
        check(globe_set_varcomp(ctx, th5), "variance components");
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
