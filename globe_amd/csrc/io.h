// File formats of the reference, for the drop-in CLIs.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace globe_io {

struct Cohort {
    int64_t nrec = 0;
    int32_t n = 0, p = 0;
    std::vector<int32_t> indiv;   // 0-based
    std::vector<double> y, times;
    std::vector<double> X;        // (p+1) columns, column-major, TIME last
    std::vector<std::string> cov_names;
};

// Whitespace table with header INDIV PHENO TIMES <covariates...>
// (checkInputs / readInputData, PCGsolver.cpp:465-587). Values are parsed as
// float, as the reference does, then widened. Throws std::runtime_error.
Cohort read_input(const std::string& path);

// Whitespace table with header INDIV PHENO <covariates...> of the reduced model
// (checkInputs / readInputData, varCompEst_reducedModel_noTime.cpp:131-248): one record
// per individual, row r is GRM row r (the reference indexes the GRM by row and never
// uses the INDIV values for V), so indiv[r] = r and times are 0; X holds every other
// column (no TIME column; p + 1 = their count, at least one). INDIV values must be
// distinct (the reference sizes its matrices by the distinct count). Throws
// std::runtime_error.
Cohort read_input_reduced(const std::string& path);

// PLINK .bed (SNP-major, magic 6C 1B 01) + .bim line count (prepGenotypeInput,
// PCGsolver.cpp:624-667), checked before the device is touched; when <prefix>.fam
// exists its line count must equal n. The rows themselves are streamed by
// globe_load_bed_file. Throws std::runtime_error.
struct BedInfo {
    int64_t m = 0;           // SNPs in .bim
    int64_t row_bytes = 0;   // ceil(n/4)
};
BedInfo check_bed(const std::string& prefix, int32_t n);

// <prefix>.grm.bin: n(n+1)/2 float32 lower triangle (readFullGRM, :778-809).
std::vector<float> read_grm(const std::string& prefix, int32_t n);

// int32 n + n float32 (saveVector, utils.cpp:5-21).
void save_vector(const std::vector<double>& v, const std::string& path);

}  // namespace globe_io
