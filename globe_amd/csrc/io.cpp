// File readers / writers of the reference's formats (host C++, used by the CLIs).
#include "io.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

namespace globe_io {

Cohort read_input(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Error opening input file: " + path);
    std::string line;
    std::getline(f, line);
    std::istringstream hs(line);
    std::string col;
    std::vector<std::string> cols;
    int ii = -1, pi = -1, ti = -1;
    while (hs >> col) {
        if (col == "INDIV") ii = (int)cols.size();
        else if (col == "PHENO") pi = (int)cols.size();
        else if (col == "TIMES") ti = (int)cols.size();
        cols.push_back(col);
    }
    if (ii < 0 || pi < 0 || ti < 0)
        throw std::runtime_error("Error: Missing required columns in header.");
    Cohort c;
    const int ncol = (int)cols.size();
    c.p = ncol - 3;
    for (int q = 0; q < ncol; ++q)
        if (q != ii && q != pi && q != ti) c.cov_names.push_back(cols[q]);
    std::vector<std::vector<float>> cov((size_t)c.p);
    std::unordered_set<int> ids;
    std::vector<float> yv, tv;
    std::vector<int> idv;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        float indiv = -1.f, pheno = 0.f, time = 0.f;
        int cq = 0;
        for (int q = 0; q < ncol; ++q) {
            float v = 0.f;
            is >> v;
            if (q == ii) indiv = v;
            else if (q == pi) pheno = v;
            else if (q == ti) time = v;
            else cov[cq++].push_back(v);
        }
        idv.push_back((int)indiv);
        ids.insert((int)indiv);
        yv.push_back(pheno);
        tv.push_back(time);
    }
    c.nrec = (int64_t)idv.size();
    c.n = (int32_t)ids.size();
    c.indiv.resize(c.nrec);
    c.y.assign(yv.begin(), yv.end());
    c.times.assign(tv.begin(), tv.end());
    c.X.resize((size_t)c.nrec * (c.p + 1));
    for (int64_t r = 0; r < c.nrec; ++r) {
        c.indiv[r] = idv[r] - 1;
        for (int q = 0; q < c.p; ++q) c.X[(size_t)q * c.nrec + r] = cov[q][r];
        c.X[(size_t)c.p * c.nrec + r] = tv[r];
    }
    return c;
}

Cohort read_input_reduced(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Error opening input file: " + path);
    std::string line;
    std::getline(f, line);
    std::istringstream hs(line);
    std::string col;
    std::vector<std::string> cols;
    int ii = -1, pi = -1;
    while (hs >> col) {
        if (col == "INDIV") ii = (int)cols.size();
        else if (col == "PHENO") pi = (int)cols.size();
        cols.push_back(col);
    }
    if (ii < 0 || pi < 0) throw std::runtime_error("Error: Missing required columns in header.");
    const int ncol = (int)cols.size();
    const int ncov = ncol - 2;
    if (ncov < 1) throw std::runtime_error("Error: at least one covariate column is required.");
    Cohort c;
    c.p = ncov - 1;
    for (int q = 0; q < ncol; ++q)
        if (q != ii && q != pi) c.cov_names.push_back(cols[q]);
    std::vector<std::vector<float>> cov((size_t)ncov);
    std::unordered_set<int> ids;
    std::vector<float> yv;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        float indiv = -1.f, pheno = 0.f;
        int cq = 0;
        for (int q = 0; q < ncol; ++q) {
            float v = 0.f;
            is >> v;
            if (q == ii) indiv = v;
            else if (q == pi) pheno = v;
            else cov[cq++].push_back(v);
        }
        if (!ids.insert((int)indiv).second)
            throw std::runtime_error("Error: individual " + std::to_string((int)indiv) +
                                     " appears twice; the reduced model takes one record per individual.");
        yv.push_back(pheno);
    }
    c.nrec = (int64_t)yv.size();
    c.n = (int32_t)c.nrec;
    c.indiv.resize(c.nrec);
    c.y.assign(yv.begin(), yv.end());
    c.times.assign(c.nrec, 0.0);
    c.X.resize((size_t)c.nrec * ncov);
    for (int64_t r = 0; r < c.nrec; ++r) {
        c.indiv[r] = (int32_t)r;
        for (int q = 0; q < ncov; ++q) c.X[(size_t)q * c.nrec + r] = cov[q][r];
    }
    return c;
}

BedInfo check_bed(const std::string& prefix, int32_t n) {
    const std::string bim = prefix + ".bim", bed = prefix + ".bed", fam = prefix + ".fam";
    BedInfo out;
    std::ifstream b(bim);
    if (!b.is_open()) throw std::runtime_error("[prepGenotypeInput] Error opening .bim file: " + bim);
    out.m = std::count(std::istreambuf_iterator<char>(b), std::istreambuf_iterator<char>(), '\n');
    std::ifstream fm(fam);
    if (fm.is_open()) {
        const int64_t nf = std::count(std::istreambuf_iterator<char>(fm), std::istreambuf_iterator<char>(), '\n');
        if (nf != n)
            throw std::runtime_error("[prepGenotypeInput] " + fam + " lists " + std::to_string(nf) +
                                     " samples but the phenotype file has N = " + std::to_string(n));
    }
    std::ifstream f(bed, std::ios::binary | std::ios::ate);
    if (!f.is_open()) throw std::runtime_error("[prepGenotypeInput] Error opening .bed file: " + bed);
    const int64_t len = (int64_t)f.tellg();
    out.row_bytes = (n + 3) / 4;
    unsigned char magic[3] = {0, 0, 0};
    f.seekg(0);
    if (len < 3 || !f.read(reinterpret_cast<char*>(magic), 3) || magic[0] != 0x6C || magic[1] != 0x1B)
        throw std::runtime_error("[prepGenotypeInput] Invalid .bed format (magic mismatch).");
    if (magic[2] == 0x00)
        throw std::runtime_error(
            "[prepGenotypeInput] BED is individual-major, code expects SNP-major => aborting.");
    if (magic[2] != 0x01) throw std::runtime_error("[prepGenotypeInput] Third byte unknown => aborting.");
    if (len < 3 + out.m * out.row_bytes)
        throw std::runtime_error("[prepGenotypeInput] .bed holds fewer than M x ceil(N/4) genotype bytes");
    return out;
}

std::vector<float> read_grm(const std::string& prefix, int32_t n) {
    const std::string path = prefix + ".grm.bin";
    const size_t nel = (size_t)n * (n + 1) / 2;
    std::vector<float> lt(nel);
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("[readFullGRM] Error opening GRM file: " + path);
    f.read(reinterpret_cast<char*>(lt.data()), (std::streamsize)(nel * sizeof(float)));
    if (!f) throw std::runtime_error("[readFullGRM] Error reading from file: " + path);
    return lt;
}

void save_vector(const std::vector<double>& v, const std::string& path) {
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) {
        std::cerr << "[saveVector] Error opening file: " << path << std::endl;
        return;
    }
    const int32_t n = (int32_t)v.size();
    std::vector<float> w(v.begin(), v.end());
    f.write(reinterpret_cast<const char*>(&n), sizeof(int32_t));
    f.write(reinterpret_cast<const char*>(w.data()), (std::streamsize)(w.size() * sizeof(float)));
    f.close();
    if (!f) std::cerr << "[saveVector] Error writing file: " << path << std::endl;
    else std::cout << "[saveVector] Vector saved: " << path << std::endl;
}

}  // namespace globe_io
