// Genotype ingest: PLINK .bed rows -> QC -> the resident 2-bit dosage layouts
// (prepGenotypeInput, PCGsolver.cpp:615-763).
//
// The rows stream through a pipeline of kSlots chunks:
//   host threads fill pinned buffer b (pread from the .bed, or memcpy from the
//   caller's rows) -> copy stream: host-to-device copy into device buffer b ->
//   compute stream: counts + FP32 keep rule (k_bed_qc), destinations of the kept rows
//   (k_keep_scan), compaction into the SNP-major layout (k_compact_pos).
// The host fill of chunk i + 1 overlaps the copy of chunk i and the QC of chunk i - 1.
// Keep decisions and destinations are made on the device, so the host never waits
// for a chunk's counts; they come back once, at the end, with the frequencies. The
// individual-major copy is then built by the transpose kernel.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <cstring>
#include <fstream>
#include <functional>
#include <iterator>
#include <thread>

#include "context.h"

namespace globe {
namespace {

constexpr int kSlots = 3;

// fill(dst, s0, cnt, threads): rows [s0, s0 + cnt) at the input pitch into pinned dst
using HostFill = std::function<void(uint8_t*, int64_t, int64_t, int)>;
// fill(dst, s0, cnt): the same rows written on the device (stream-ordered)
using DevFill = std::function<void(uint8_t*, int64_t, int64_t)>;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

// f(t, T) for t in [0, T) on T threads; the calling thread runs t = 0.
void parallel(int T, const std::function<void(int, int)>& f) {
    std::vector<std::thread> th;
    th.reserve(T > 1 ? T - 1 : 0);
    for (int t = 1; t < T; ++t) th.emplace_back(f, t, T);
    f(0, T);
    for (auto& x : th) x.join();
}

// Pinned buffers, the copy stream and its events; the destructor drains both streams
// before anything it owns (or the device buffers declared before it) is freed.
struct Pipeline {
    hipStream_t compute = nullptr, copy = nullptr;
    void* pin[kSlots] = {nullptr, nullptr, nullptr};
    hipEvent_t h2d[kSlots] = {nullptr, nullptr, nullptr}, done[kSlots] = {nullptr, nullptr, nullptr};
    int slots = 0;
    Pipeline(hipStream_t s, int nslots, size_t bytes) : compute(s), slots(nslots) {
        HIP_CHECK(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
        for (int b = 0; b < slots; ++b) {
            HIP_CHECK(hipHostMalloc(&pin[b], bytes, hipHostMallocDefault));
            HIP_CHECK(hipEventCreateWithFlags(&h2d[b], hipEventDisableTiming));
            HIP_CHECK(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
        }
    }
    ~Pipeline() {
        if (copy) (void)hipStreamSynchronize(copy);
        (void)hipStreamSynchronize(compute);
        for (int b = 0; b < kSlots; ++b) {
            if (pin[b]) (void)hipHostFree(pin[b]);
            if (h2d[b]) (void)hipEventDestroy(h2d[b]);
            if (done[b]) (void)hipEventDestroy(done[b]);
        }
        if (copy) (void)hipStreamDestroy(copy);
    }
};

// Stage m input SNP rows of pitch_in bytes from a host source (hf) or a device
// generator (df): QC per row, kept rows compacted in SNP order, both layouts built.
void stage_rows(globe_ctx* c, int64_t m, int64_t pitch_in, float min_maf, float max_miss,
                const HostFill* hf, const DevFill* df) {
    require(c->has_cohort, GLOBE_E_STATE, "cohort must be set before genotypes");
    const int32_t n = c->n;
    require(pitch_in >= (n + 3) / 4, GLOBE_E_INVALID, "row pitch smaller than ceil(N/4)");
    require(min_maf >= 0.0f && min_maf <= 0.5f && max_miss >= 0.0f && max_miss <= 1.0f,
            GLOBE_E_INVALID, "invalid MAF or missing rate thresholds");
    const double t0 = now_s();
    hipStream_t s = c->stream;
    c->gen = 0;
    c->gdiag_ok = false;
    c->d_grm.release();
    c->pitch_s = packed_pitch(n);
    c->d_snp.release();
    c->d_ind.release();
    // rows are padded to a multiple of kRowPad with zero (dosage-0) bytes so the
    // contraction kernels load whole row blocks without bounds checks
    c->d_snp.ensure((size_t)packed_rows(m) * c->pitch_s);
    // the synthetic generator seeds each 1 GiB chunk of rows, so its chunking is fixed
    const int64_t target = hf ? c->ingest_chunk : (1ll << 30);
    const int64_t chunk_rows = std::max<int64_t>(1, std::min<int64_t>(m, target / pitch_in));
    const int slots = hf ? kSlots : 1;
    const size_t chunk_bytes = (size_t)chunk_rows * pitch_in;
    DevBuf raw[kSlots], d_pos, d_alt, d_miss, d_maf, d_keep, d_base;
    for (int b = 0; b < slots; ++b) raw[b].ensure(chunk_bytes);
    const size_t mm = (size_t)std::max<int64_t>(m, 1);
    d_pos.ensure(sizeof(int64_t) * (size_t)chunk_rows * slots);
    d_alt.ensure(sizeof(int64_t) * mm);
    d_miss.ensure(sizeof(int64_t) * mm);
    d_maf.ensure(sizeof(float) * mm);
    d_keep.ensure(mm);
    d_base.ensure(sizeof(int64_t));
    HIP_CHECK(hipMemsetAsync(d_base.p, 0, sizeof(int64_t), s));
    Pipeline pl(s, hf ? slots : 0, chunk_bytes);
    const int T = host_threads();
    double t_fill = 0, t_wait = 0;
    int64_t chunks = 0;
    for (int64_t s0 = 0; s0 < m; s0 += chunk_rows, ++chunks) {
        const int b = (int)(chunks % slots);
        const int64_t cnt = std::min(chunk_rows, m - s0);
        uint8_t* dev = raw[b].as<uint8_t>();
        int64_t* pos = d_pos.as<int64_t>() + (size_t)b * chunk_rows;
        if (hf) {
            uint8_t* pin = static_cast<uint8_t*>(pl.pin[b]);
            if (chunks >= slots) {  // the pinned buffer's previous copy must have landed
                const double tw = now_s();
                HIP_CHECK(hipEventSynchronize(pl.h2d[b]));
                t_wait += now_s() - tw;
            }
            const double tf = now_s();
            (*hf)(pin, s0, cnt, T);
            t_fill += now_s() - tf;
            // the device buffer is free once the compaction that read it has run
            if (chunks >= slots) HIP_CHECK(hipStreamWaitEvent(pl.copy, pl.done[b], 0));
            HIP_CHECK(hipMemcpyAsync(dev, pin, (size_t)cnt * pitch_in, hipMemcpyHostToDevice, pl.copy));
            HIP_CHECK(hipEventRecord(pl.h2d[b], pl.copy));
            HIP_CHECK(hipStreamWaitEvent(s, pl.h2d[b], 0));
        } else {
            (*df)(dev, s0, cnt);
        }
        launch_bed_qc(s, dev, cnt, pitch_in, n, min_maf, max_miss, d_alt.as<int64_t>() + s0,
                      d_miss.as<int64_t>() + s0, d_maf.as<float>() + s0, d_keep.as<uint8_t>() + s0);
        launch_keep_scan(s, d_keep.as<uint8_t>() + s0, cnt, d_base.as<int64_t>(), pos);
        launch_compact_pos(s, dev, pitch_in, pos, cnt, n, c->d_snp.as<uint8_t>(), c->pitch_s);
        if (hf) HIP_CHECK(hipEventRecord(pl.done[b], s));
    }
    c->qc_alt.assign(m, 0);
    c->qc_miss.assign(m, 0);
    c->qc_keep.assign(m, 0);
    c->qc_maf.assign(m, 0.f);
    int64_t kept = 0;
    if (m > 0) {
        HIP_CHECK(hipMemcpyAsync(c->qc_alt.data(), d_alt.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(c->qc_miss.data(), d_miss.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(c->qc_maf.data(), d_maf.p, sizeof(float) * m, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(c->qc_keep.data(), d_keep.p, (size_t)m, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(&kept, d_base.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    c->h_freq.clear();
    c->h_freq.reserve((size_t)kept);
    for (int64_t q = 0; q < m; ++q)
        if (c->qc_keep[q]) c->h_freq.push_back(c->qc_maf[q]);
    require((int64_t)c->h_freq.size() == kept, GLOBE_E_DEVICE, "kept-row count mismatch after QC");
    for (int b = 0; b < kSlots; ++b) raw[b].release();
    c->m_total = m;
    c->m = kept;
    {
        const int64_t pad_rows = packed_rows(kept) - kept;
        HIP_CHECK(hipMemsetAsync(c->d_snp.as<uint8_t>() + kept * c->pitch_s, 0,
                                 (size_t)pad_rows * c->pitch_s, s));
    }
    c->pitch_i = packed_pitch(kept);
    c->d_ind.ensure((size_t)packed_rows(n) * c->pitch_i);
    HIP_CHECK(hipMemsetAsync(c->d_ind.p, 0, (size_t)packed_rows(n) * c->pitch_i, s));
    launch_transpose_geno(s, c->d_snp.as<uint8_t>(), kept, c->pitch_s, n, c->d_ind.as<uint8_t>(), c->pitch_i);
    c->d_freq.ensure(sizeof(float) * std::max<int64_t>(kept, 1));
    c->d_twof.ensure(sizeof(double) * std::max<int64_t>(kept, 1));
    c->d_inv.ensure(sizeof(double) * std::max<int64_t>(kept, 1));
    if (kept > 0) {
        HIP_CHECK(hipMemcpyAsync(c->d_freq.p, c->h_freq.data(), sizeof(float) * kept, hipMemcpyHostToDevice, s));
        launch_snp_consts(s, c->d_freq.as<float>(), kept, c->d_twof.as<double>(), c->d_inv.as<double>());
    }
    HIP_CHECK(hipStreamSynchronize(s));
    c->gen = 1;
    c->gen_n = n;
    c->ingest_stats[0] = now_s() - t0;
    c->ingest_stats[1] = t_fill;
    c->ingest_stats[2] = t_wait;
    c->ingest_stats[3] = (double)chunks;
    c->ingest_stats[4] = hf ? (double)T : 0.0;
}

// Lines in a text file (the reference counts '\n' in the .bim, PCGsolver.cpp:634-636).
int64_t count_lines(std::ifstream& f) {
    return (int64_t)std::count(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>(), '\n');
}

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

}  // namespace
}  // namespace globe

using namespace globe;

extern "C" {

int globe_load_bed(globe_ctx* c, const uint8_t* rows, int64_t m, int64_t pitch, float min_maf,
                   float max_miss, int64_t* m_kept) {
    API_BEGIN(c)
    require(rows != nullptr || m == 0, GLOBE_E_INVALID, "null genotype rows");
    require(m >= 0, GLOBE_E_INVALID, "negative SNP count");
    HostFill fill = [&](uint8_t* dst, int64_t s0, int64_t cnt, int T) {
        const size_t bytes = (size_t)cnt * pitch;
        const uint8_t* src = rows + (size_t)s0 * pitch;
        const int TT = bytes >= (8u << 20) ? T : 1;
        parallel(TT, [&](int t, int NT) {
            const size_t lo = bytes * t / NT, hi = bytes * (t + 1) / NT;
            std::memcpy(dst + lo, src + lo, hi - lo);
        });
    };
    stage_rows(c, m, pitch, min_maf, max_miss, &fill, nullptr);
    if (m_kept) *m_kept = c->m;
    API_END(c)
}

int globe_load_bed_file(globe_ctx* c, const char* prefix, float min_maf, float max_miss,
                        int64_t* m_total, int64_t* m_kept) {
    API_BEGIN(c)
    require(prefix != nullptr, GLOBE_E_INVALID, "null genotype prefix");
    require(c->has_cohort, GLOBE_E_STATE, "cohort must be set before genotypes");
    const std::string pre(prefix), bim = pre + ".bim", bed = pre + ".bed", fam = pre + ".fam";
    const int32_t n = c->n;
    int64_t m = 0;
    {
        std::ifstream f(bim);
        require(f.is_open(), GLOBE_E_IO, "[prepGenotypeInput] Error opening .bim file: " + bim);
        m = count_lines(f);
    }
    {
        // the reference never opens the .fam (PCGsolver.cpp:626); when present, its
        // sample count must be the cohort's, or every genotype would be misassigned
        std::ifstream f(fam);
        if (f.is_open()) {
            const int64_t nf = count_lines(f);
            require(nf == n, GLOBE_E_INVALID,
                    "[prepGenotypeInput] " + fam + " lists " + std::to_string(nf) +
                        " samples but the phenotype file has N = " + std::to_string(n));
        }
    }
    Fd fd;
    fd.fd = ::open(bed.c_str(), O_RDONLY);
    require(fd.fd >= 0, GLOBE_E_IO, "[prepGenotypeInput] Error opening .bed file: " + bed);
    struct stat st;
    require(fstat(fd.fd, &st) == 0, GLOBE_E_IO, "[prepGenotypeInput] cannot stat " + bed);
    uint8_t magic[3] = {0, 0, 0};
    require(st.st_size >= 3 && ::pread(fd.fd, magic, 3, 0) == 3, GLOBE_E_INVALID,
            "[prepGenotypeInput] Invalid .bed format (magic mismatch).");
    require(magic[0] == 0x6C && magic[1] == 0x1B, GLOBE_E_INVALID,
            "[prepGenotypeInput] Invalid .bed format (magic mismatch).");
    require(magic[2] != 0x00, GLOBE_E_INVALID,
            "[prepGenotypeInput] BED is individual-major, code expects SNP-major => aborting.");
    require(magic[2] == 0x01, GLOBE_E_INVALID, "[prepGenotypeInput] Third byte unknown => aborting.");
    const int64_t row_bytes = (n + 3) / 4;
    require((int64_t)st.st_size >= 3 + m * row_bytes, GLOBE_E_INVALID,
            "[prepGenotypeInput] .bed holds fewer than M x ceil(N/4) genotype bytes");
    std::atomic<int> io_err{0};
    HostFill fill = [&](uint8_t* dst, int64_t s0, int64_t cnt, int T) {
        const size_t bytes = (size_t)cnt * row_bytes;
        const off_t base = 3 + (off_t)s0 * row_bytes;
        const int TT = bytes >= (8u << 20) ? T : 1;
        parallel(TT, [&](int t, int NT) {
            size_t lo = bytes * t / NT;
            const size_t hi = bytes * (t + 1) / NT;
            while (lo < hi) {
                const ssize_t r = ::pread(fd.fd, dst + lo, hi - lo, base + (off_t)lo);
                if (r < 0 && errno == EINTR) continue;
                if (r <= 0) {
                    io_err.store(r == 0 ? -1 : errno);
                    return;
                }
                lo += (size_t)r;
            }
        });
        const int e = io_err.load();
        require(e == 0, GLOBE_E_IO,
                "[prepGenotypeInput] read error in " + bed + (e > 0 ? std::string(": ") + std::strerror(e) : ""));
    };
    stage_rows(c, m, row_bytes, min_maf, max_miss, &fill, nullptr);
    if (m_total) *m_total = m;
    if (m_kept) *m_kept = c->m;
    API_END(c)
}

int globe_synth_bed(globe_ctx* c, int64_t m, uint64_t seed, float f_lo, float f_hi, float miss,
                    float min_maf, float max_miss, int64_t* m_kept) {
    API_BEGIN(c)
    require(m >= 0 && f_lo >= 0.f && f_hi <= 1.f && f_lo <= f_hi && miss >= 0.f && miss <= 1.f,
            GLOBE_E_INVALID, "globe_synth_bed: bad parameters");
    require(c->has_cohort, GLOBE_E_STATE, "cohort must be set before genotypes");
    const int64_t pitch = packed_pitch(c->n);
    DevFill gen = [&](uint8_t* dst, int64_t s0, int64_t cnt) {
        launch_synth_bed(c->stream, dst, cnt, pitch, c->n, seed + 0x51ED2701ull * (uint64_t)s0, f_lo, f_hi,
                         miss);
    };
    stage_rows(c, m, pitch, min_maf, max_miss, nullptr, &gen);
    if (m_kept) *m_kept = c->m;
    API_END(c)
}

int globe_set_ingest_chunk(globe_ctx* c, int64_t bytes) {
    API_BEGIN(c)
    require(bytes >= 1, GLOBE_E_INVALID, "ingest chunk must be at least one byte");
    c->ingest_chunk = bytes;
    API_END(c)
}

int globe_ingest_stats(globe_ctx* c, double out[5]) {
    API_BEGIN(c)
    require(out != nullptr, GLOBE_E_INVALID, "null output");
    for (int q = 0; q < 5; ++q) out[q] = c->ingest_stats[q];
    API_END(c)
}

}  // extern "C"
