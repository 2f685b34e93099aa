// Internal declarations of libglobe_amd.so: device buffers, the context, and the
// host-side launch wrappers of the HIP kernels (k_record.hip, k_geno.hip,
// k_grm.hip, k_mfma.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/globe.h"

namespace globe {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
// A failed check throws Error(code, msg). The literal form builds no string unless the
// check fails (checks sit in per-record loops of request and input validation).
inline void require(bool cond, int code, const char* msg) {
    if (!cond) throw Error(code, msg);
}
inline void require(bool cond, int code, const std::string& msg) {
    if (!cond) throw Error(code, msg);
}
#define HIP_CHECK(x) ::globe::hip_check((x), #x)

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b);
    void release();
    ~DevBuf() { release(); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

constexpr int kBlock = 256;         // threads per block of the streaming kernels
constexpr int kMaxParts = 1024;     // partial-sum slots per column (grid cap)
constexpr int kMaxCols = 64;        // max right-hand sides per solve
constexpr int kGenoTile = 512;      // genotypes per K-step of the row-dot kernels
constexpr int64_t kPitchAlign = 128;  // bytes: row pitch multiple (512 genotypes)
constexpr int64_t kRowPad = 512;      // packed layouts hold a multiple of this many rows

inline int parts_for(int64_t count) {
    int64_t b = (count + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    return (int)(b > kMaxParts ? kMaxParts : b);
}
inline int64_t round_up(int64_t a, int64_t m) { return (a + m - 1) / m * m; }
// Packed 2-bit layouts: bytes per row holding `count` genotypes, and rows allocated for
// `rows` rows (zero rows past the last one, so row blocks read without bounds checks)
inline int64_t packed_pitch(int64_t count) { return round_up((count + 3) / 4, kPitchAlign); }
inline int64_t packed_rows(int64_t rows) { return round_up(rows > 1 ? rows : 1, kRowPad); }

// Device-side PCG state: scalars then per-column arrays (k <= kMaxCols).
struct PcgState {
    int done, iters, stag, stalled;
    int max_iter, rule, stagnation, k;
    double tol, last, max_ratio, pad;
    double rs[kMaxCols], alpha[kMaxCols], beta[kMaxCols], bnorm[kMaxCols], ratio[kMaxCols];
    int active[kMaxCols], col_iters[kMaxCols];
    // Lanczos coefficients of the solve (nullptr: not kept): hist[c] = r0.z0 of column
    // c, then per iteration it < hist_cap the row pair hist[(1 + 2 it) kMaxCols + c] =
    // alpha_it and hist[(2 + 2 it) kMaxCols + c] = beta_it of the columns still active
    double* hist;
    int hist_cap, hist_pad;
};

// Digit layout of a k-column right-hand side in the int8 contraction: column c has L[c]
// int8 digits (0..7; 0: a column whose product is not needed, computed as zero) in digit
// columns off[c] .. off[c] + L[c] - 1; digit column j holds digit dig[j] of column col[j]
// (col = 255: padding). Passed to kernels by value.
struct LimbMap {
    int8_t L[kMaxCols];
    uint8_t off[kMaxCols];
    uint8_t col[128];
    uint8_t dig[128];
};

// ---------------------------------------------------------------- record space
// Record-space structure seen by the record kernels. The records of an individual
// are contiguous (a block); column c of a block PCG may carry its own structure (the
// serving slots) at column strides rec_ld / blk_ld / tile_ld, or all columns share one
// (strides 0, the cohort). Tiles are runs of whole blocks of at most `cap` records
// (tile t = blocks tile_blk[t] .. tile_blk[t+1]-1; rec_blk[r] = block of record r),
// staged through LDS by the tiled kernels; cap == 0 selects the thread-per-block
// kernels (shared structure only: an individual with more than kTileMaxRecords records).
struct RecMeta {
    const int32_t* indiv = nullptr;    // individual of each record
    const double* times = nullptr;
    const int32_t* rec_blk = nullptr;
    int64_t rec_ld = 0;
    const int64_t* blk_off = nullptr;  // nblk + 1 record offsets
    const int32_t* blk_ind = nullptr;  // individual of each block
    int64_t blk_ld = 0;
    const int32_t* tile_blk = nullptr;
    int64_t tile_ld = 0;
    const int32_t* ntiles_col = nullptr;  // per-column tile counts (device) or null
    int32_t ntiles = 0;                   // tiles (max over columns when ntiles_col is set)
    const int64_t* nrec_col = nullptr;    // per-column record counts (device) or null
    // per-column start (device, elements) of every record-space array -- structure and
    // vectors -- when the columns live at their own offsets of one pool (the serving
    // loop); null: column c starts at c times the stride of each array
    const int64_t* col_off = nullptr;
    int64_t nrec = 0;                     // records (row count of the record matrices)
    int32_t nblk = 0;                     // blocks (shared structure)
    int32_t cap = 0;                      // records per tile; 0: thread-per-block kernels
};
constexpr int kTileRecords = 2048;
constexpr int kTileMaxRecords = 2048;  // LDS: 2 doubles per record slot (32 KB)
// individuals (blocks) per tile: the preconditioner keeps 2 doubles per block in LDS
// (8 KB), so its footprint is 40 KB and 4 workgroups fit a CU
constexpr int kTileBlocks = 512;
// Tiles end at kTileRecords records or kTileBlocks blocks, so two consecutive tiles hold
// at least kTileBlocks records: a column of R records has at most this many tiles.
inline int64_t max_tiles(int64_t records) { return 2 * records / kTileBlocks + 1; }
// partial-sum slots written by launch_precond
int precond_parts(const RecMeta& m);
void launch_seg_sums(hipStream_t s, const RecMeta& m, const double* x, int64_t ldx, int k,
                     double* S0, double* S1, int64_t ldn, const int* skip);
// PCG forms on tiled record structure (return false, launching nothing, on the
// thread-per-block structure): the segmented sums of the search direction after its
// update p = z + beta p (active columns, skipped once the solve is done), and alpha from
// the np p.Vp partials with the x/r update and the preconditioner (r.z and r.r partials
// of the new r, in precond_parts(m) slots)
bool launch_seg_sums_pcg(hipStream_t s, const RecMeta& m, double* p, const double* z, int64_t ld, int k,
                         PcgState* st, double* S0, double* S1, int64_t ldn);
bool launch_pcg_xr_precond(hipStream_t s, const RecMeta& m, const double* gdiag, const double* theta_cols,
                           PcgState* st, const double* part_pvp, int np, double* x, int64_t ldx, double* r,
                           const double* p, const double* vp, int64_t ld, int k, double* z, double* part_rz,
                           double* part_rr);
void launch_record_out(hipStream_t s, const RecMeta& m, const double* x, int64_t ldx, int k,
                       const double* GZ, const double* S0, const double* S1, int64_t ldn,
                       const double* theta_cols, double* y, int64_t ldy, double* part_dot,
                       const int* skip);
// st: inside PCG (frozen columns skipped, nothing when done)
void launch_precond(hipStream_t s, const RecMeta& m, const double* gdiag, const double* theta_cols,
                    const double* r, int64_t ldr, int k, double* z, int64_t ldz, double* part_rz,
                    double* part_rr, const PcgState* st, int tilde_v);
void launch_dot_parts(hipStream_t s, int64_t nrec, const double* a, int64_t lda, const double* b,
                      int64_t ldb, int k, double* part);
void launch_reduce_parts(hipStream_t s, const double* part, int nparts, int k, double* out);
void launch_gram(hipStream_t s, int64_t nrec, const double* A, int64_t lda, int ka,
                 const double* B, int64_t ldb, int kb, double* part, double* out);
// out[a kb + b] = A[:, a] . B[:, a kb + b] for a < ka, b < kb (each column of A against its
// own group of kb columns of B); part: kMaxParts x ka kb
void launch_gram_grouped(hipStream_t s, int64_t nrec, const double* A, int64_t lda, int ka,
                         const double* B, int64_t ldb, int kb, double* part, double* out);
void launch_axpy_cols(hipStream_t s, int64_t nrec, double* y, int64_t ldy, const double* x,
                      int64_t ldx, int kx, const double* coef, int ky);
// pooled column j: offset, tile and record counts, components (written on the device)
void launch_set_column(hipStream_t s, int j, int64_t off, int32_t ntiles, int64_t nrec, const double theta[5],
                       int64_t* col_off, int32_t* ntiles_col, int64_t* nrec_col, double* theta_cols);
void launch_pcg_init(hipStream_t s, PcgState* st, const double* part_bb, const double* part_rr,
                     const double* part_rz, int nparts, int k);
void launch_pcg_admit(hipStream_t s, PcgState* st, int j, const double* part_bb, int nparts_bb,
                      const double* part_rr, const double* part_rz, int nparts);
// one-block reductions of the dot partials: alpha, the convergence test, beta
void launch_pcg_alpha(hipStream_t s, PcgState* st, const double* part_pvp, int nparts, int k);
void launch_pcg_check(hipStream_t s, PcgState* st, const double* part_rr, int nparts, int k);
void launch_pcg_beta(hipStream_t s, PcgState* st, const double* part_rz, int nparts, int k);
// column j of a block PCG stops iterating
void launch_pcg_deactivate(hipStream_t s, PcgState* st, int j);
// the convergence test, then beta unless the solve has ended
void launch_pcg_check_beta(hipStream_t s, PcgState* st, const double* part_rr, int nparts_rr,
                           const double* part_rz, int nparts_rz, int k);
// x += alpha p, r -= alpha Vp per active column with the r.r partials
void launch_pcg_xr(hipStream_t s, PcgState* st, int64_t nrec, double* x, int64_t ldx, double* r,
                   const double* p, const double* vp, int64_t ld, int k, double* part_rr,
                   const int64_t* col_off = nullptr, const int64_t* col_cnt = nullptr);
// col_off / col_cnt (device, may be null): column starts and record counts of pooled
// columns (otherwise column c starts at c ld and holds nrec records)
void launch_pcg_p(hipStream_t s, PcgState* st, int64_t nrec, double* p, const double* z,
                  int64_t ld, int k, const int64_t* col_off = nullptr, const int64_t* col_cnt = nullptr);
void launch_bhat(hipStream_t s, int32_t nblk, const int64_t* blk_off, const int32_t* blk_ind,
                 const double* times, const double* v, const double* theta, double* bhat);
void launch_f_theta(hipStream_t s, int32_t nblk, const int64_t* blk_off, const int32_t* blk_ind,
                    const double* times, const double* py, int64_t ldpy, const double* uhat,
                    int64_t ldu, int ncols, double inv_sg, double* f, int64_t ldf);
void launch_copy_cols(hipStream_t s, int64_t rows, const double* src, int64_t lds, double* dst,
                      int64_t ldd, int k);
void launch_fill(hipStream_t s, double* dst, int64_t count, double v);
// zero partial slots [from, to) of k columns of up to three partial arrays (null: skipped)
void launch_zero_slots(hipStream_t s, double* a, double* b, double* c, int k, int from, int to);
void launch_copy16(hipStream_t s, void* dst, const void* src, int64_t bytes);
void launch_sub_cols(hipStream_t s, int64_t rows, const double* a, int64_t lda, const double* b,
                     int64_t ldb, double* d, int64_t ldd, int k);
void launch_fill_cols(hipStream_t s, int64_t rows, double* d, int64_t ldd, int k, double v);

// ---------------------------------------------------------------- genotypes
// counts + keep rule per row (maf/keep may be null: counts only)
void launch_bed_qc(hipStream_t s, const uint8_t* rows, int64_t m, int64_t pitch, int32_t n,
                   float min_maf, float max_miss, int64_t* alt, int64_t* miss, float* maf,
                   uint8_t* keep);
// pos[q] = destination row of kept row q (or -1); *base advances by the kept count
void launch_keep_scan(hipStream_t s, const uint8_t* keep, int64_t cnt, int64_t* base, int64_t* pos);
void launch_compact_pos(hipStream_t s, const uint8_t* src, int64_t pitch_src, const int64_t* pos,
                        int64_t cnt, int32_t n, uint8_t* dst, int64_t pitch_dst);
void launch_transpose_geno(hipStream_t s, const uint8_t* snp, int64_t m, int64_t pitch_s, int32_t n,
                           uint8_t* ind, int64_t pitch_i);
void launch_synth_bed(hipStream_t s, uint8_t* rows, int64_t m, int64_t pitch, int32_t n,
                      uint64_t seed, float f_lo, float f_hi, float miss);
void launch_snp_consts(hipStream_t s, const float* freq, int64_t m, double* twof, double* inv);
void launch_gdiag(hipStream_t s, const uint8_t* ind, int64_t pitch_i, int32_t n, int64_t m,
                  const double* twof, const double* inv, double* gdiag);
// out[c][r] = sum_k dosage(rows[r], k) * v[c][k] over nk genotypes (FP64).
void launch_rowdot(hipStream_t s, const uint8_t* rows, int64_t pitch, int64_t nrows, int64_t nk,
                   const double* v, int64_t ldv, int k, double* out, int64_t ldo, const int* skip);
void launch_center_snp(hipStream_t s, int64_t m, int k, const double* P1, int64_t ld1,
                       const double* T, const double* twof, const double* inv,
                       const double* theta_cols, double inv_m, double* u, int64_t ldu,
                       double* part_c, const int* skip);
void launch_center_ind(hipStream_t s, int32_t n, int k, const double* P2, int64_t ld2,
                       const double* Csum, double* GZ, int64_t ldg, const int* skip);
void launch_colsum_parts(hipStream_t s, int64_t rows, const double* v, int64_t ldv, int k,
                         double* part, const int* skip);
void launch_reduce_parts_skip(hipStream_t s, const double* part, int nparts, int k, double* out,
                              const int* skip);

// ---------------------------------------------------------------- probes (k_probe.hip)
// nmc columns (leading dimension ldy) of Philox-seeded probes: kind 0 Gaussian tilde-V at
// theta (per-record individual and time), 1 Rademacher, 2 unit vectors sqrt(nrec) e_s
// log-determinant probes b = F u, F F^T = Q (k_probe.hip); probe s is number first + s
void launch_logdet_probes(hipStream_t s, int kind, int64_t nrec, int64_t first, int count, uint64_t seed,
                          const int32_t* indiv, const double* times, const double* gdiag,
                          const double theta[5], double* Y, int64_t ldy);
// partial sums of log|Q_i| over the blocks into part; returns the partial count
int launch_logdet_q(hipStream_t s, int32_t nblk, const int64_t* off, const int32_t* ind, const double* t,
                    const double* gdiag, const double* theta, double* part);
void launch_probes(hipStream_t s, int kind, int64_t nrec, int nmc, uint64_t seed, const int32_t* indiv,
                   const double* times, const double theta[5], double* Y, int64_t ldy);

// ---------------------------------------------------------------- GRM
// The device GRM is tile-major (k_grm.hip): 64 x 64 FP32 tiles (I, J <= I) at tile
// index I(I+1)/2 + J, diagonal tiles mirrored, zero past n; grm_tile_bytes(n) bytes.
size_t grm_tile_bytes(int32_t n);
// packed-triangle rows r0 .. r1 - 1 (lt_rows starts at row r0) <-> tiles
void launch_grm_tiles_from_lt(hipStream_t s, const float* lt_rows, int64_t r0, int64_t r1, float* tiles);
void launch_grm_lt_from_tiles(hipStream_t s, const float* tiles, int64_t r0, int64_t r1, float* lt_rows);
// part: grm_product_part_bytes(n) of scratch for the strip partials and S0 slices
// k <= kSymMaxCols columns: every stored tile read once (scratch: grm_sym_part_bytes(n, k))
constexpr int kSymMaxCols = 8;
size_t grm_sym_part_bytes(int32_t n, int k);
void launch_grm_product_sym(hipStream_t s, const float* tiles, int32_t n, const double* S0, int64_t ldn, int k,
                            const double* theta_cols, double* GZ, int64_t ldg, const int* skip, double* scratch);
void launch_grm_product(hipStream_t s, const float* tiles, int32_t n, const double* S0, int64_t ldn,
                        int k, const double* theta_cols, double* GZ, int64_t ldg, const int* skip,
                        double* part);
size_t grm_product_part_bytes(int32_t n);
// columns the FP64 strip product computes for k columns (16, 20, 24 or 32 per pass of 32)
int grm_padded_cols(int k);

// Mixed-precision GRM product (k_grm_i8.hip): int8 digit tiles of the full square with
// per-row exponents, built from the FP32 tiles (aux: grm_digit_aux_bytes of scratch that
// keeps the row exponents, read back by grm_digit_rexp).
size_t grm_digit_bytes(int32_t n);
size_t grm_digit_aux_bytes(int32_t n);
void launch_grm_digits(hipStream_t s, const float* tiles, int32_t n, void* aux, int8_t* G8);
const int32_t* grm_digit_rexp(const void* aux, int32_t n);
// GZ = sig2g_c (G S0) with L (2..4) digits of G and S0; gdiag: the FP64 diagonal of G;
// scratch: grm_i8_part_bytes(n)
size_t grm_i8_part_bytes(int32_t n);
void launch_grm_product_i8(hipStream_t s, const int8_t* G8, const int32_t* rexp, int32_t n, const double* S0,
                           int64_t ldn, int k, int L, const double* gdiag, const double* theta_cols, double* GZ,
                           int64_t ldg, const int* skip, void* scratch);
// tile column j0 / 64 of G (FP32) for i >= j0 + c from the n x k columns Gc (j0 % 64 == 0)
void launch_grm_scatter(hipStream_t s, const double* Gc, int64_t ldg, int32_t n, int64_t j0, int k,
                        float* tiles);
void launch_grm_diag(hipStream_t s, const float* tiles, int32_t n, double* d);

}  // namespace globe
