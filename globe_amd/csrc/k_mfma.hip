// Packed-genotype contraction on the int8 matrix cores.
//
//   Out[r][c] = sum_kappa dosage(r, kappa) * Z[c][kappa]
//
// The dosages 0/1/2 are exact int8. Each column of Z is written in block fixed
// point, q = round(z * 2^sh_c) with |q| < 2^(8L-2) (sh_c from the column's max |z|),
// and q is split into L balanced base-256 int8 digits. Every MFMA accumulates
// exactly in int32 (|acc| <= 256 K < 2^31 for K < 8.3M genotypes), so the product
// is exact for the quantized input; the only rounding is the quantization
// (2^-(8L-2) of the column max) and the final FP64 combination of the L digit sums.
//   L = 7: 54-bit fixed point, as accurate as an FP64 dot product (GLOBE_PREC_FP64)
//   L = 4: 30-bit fixed point (GLOBE_PREC_MIXED); 4 int8 digits cost the MFMA time of
//          two BF16 pieces and keep 30 bits with exact accumulation.
// Centering and scaling (w = (g - 2f) inv) are applied afterwards as the FP64
// rank-one correction sum_i w_is z_i = inv_s (P_s - 2 f_s T), T = sum_i z_i of the
// quantized z, so the operator is applied exactly to the quantized vector.
//
// The packed rows hold 2-bit dosages (k_geno.hip), so a dword of 16 genotypes
// becomes 16 int8 dosages with one shift and one mask per 4 genotypes, in the order
// {0,4,8,12, 1,5,9,13, 2,6,10,14, 3,7,11,15}; the digit matrix Bq holds each lane's
// 16 digits in exactly that order.
//
// Two kernels:
//   thin (digit columns <= 16, the 1-RHS solve): v_mfma_i32_16x16x64_i8, W waves of RT
//     row fragments of 16 rows. A stage is one super-step of 256 genotypes = 64 packed
//     bytes of a row; A lane l: row l % 16, packed bytes 64 ss + 16 (l / 16) + 4 s for
//     K-step s < 4.
//   wide (digit columns 17..128): v_mfma_i32_32x32x32_i8, W waves of 2 row fragments
//     of 32 rows and JT column fragments of 32 (64 W rows per block; W = 8 with one
//     block per CU or W = 4 with two). A stage is 128 genotypes = 32 packed bytes of a
//     row: A lane l: row l % 32, packed bytes 32 H + 16 (l / 32) + 4 s for K-step
//     s < 4 of stage H. Both operands of a stage (2 W KB of packed rows, 4 JT KB of
//     digits) are copied global -> LDS by the buffer-to-LDS path
//     (global_load_lds_dwordx4: 64 lanes x 16 B land as one contiguous KB) into a
//     ring of NS slots, NS - 1 stages ahead of the one being multiplied; one wait +
//     barrier per stage.
// Both kernels stage every operand through LDS this way: no loaded value lives in
// VGPRs across a stage. Copies into LDS and plain VGPR loads share vmcnt, so a
// compiler-generated wait on a VGPR load would have to be vmcnt(0) and drain the ring;
// with LDS-only operands the explicit per-stage waits are the only ones.
// Output fragments (probed on gfx950): 16x16: element v of lane l is (row 4 (l/16) + v,
// column l % 16); 32x32: (row 8 (v/4) + 4 (l/32) + v % 4, column l % 32).
//
// Split-K is a power of two S; block b works on K-split b % S of row block b / S for
// S <= 8, and for S > 8 every XCD (b % 8) walks its splits one after another
// (split = b % 8 + 8 ((b / 8) / R)), so the blocks resident on one XCD share one
// K-chunk of Bq in that XCD's L2. Partial sums per split go to out[split][row][J16]
// and are added exactly (int64) in the epilogue.
#include <algorithm>
#include <tuple>
#include <mutex>
#include <map>
#include <cmath>
#include <cstring>
#include <vector>
#include <cstdio>
#include <cstdlib>

#include "context.h"
#include "contract_plan.h"

namespace globe {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr uint32_t kFieldMask = 0x03030303u;

__device__ inline v4i decode16(uint32_t w) {
    // byte j of (w >> 2d) & 0x03030303 is the dosage of genotype 4 j + d
    v4i d;
    d[0] = (int)(w & kFieldMask);
    d[1] = (int)((w >> 2) & kFieldMask);
    d[2] = (int)((w >> 4) & kFieldMask);
    d[3] = (int)((w >> 6) & kFieldMask);
    return d;
}

__device__ inline uint32_t dword_of(const uint4& v, int s) {
    return s == 0 ? v.x : s == 1 ? v.y : s == 2 ? v.z : v.w;
}

using plan::ThinCfg;
using plan::WideCfg;

// ---------------------------------------------------------------- thin kernel
// W waves x RT row fragments of 16 rows (16 W RT rows per block), one column fragment
// of 16 digits. A stage is SS super-steps (256 SS genotypes = 64 SS packed bytes of a
// row): for each of the wave's RT row fragments, SS copies (lane l: row l % 16, the 16
// bytes at 64 h + 16 (l / 16) of super-step h), and the stage's 4 SS digit chunks of
// 1 KB (spread over the waves; with more waves than chunks, one each for the first
// waves) go global -> LDS
// into a ring of NS slots, NS - 1 stages ahead; one wait + barrier per stage (the digit
// chunks are shared by the block). With SS = 2 the two 64-byte halves of every 128-byte
// line of a row are requested back to back: measured on MI355X, 16-row fragments read
// 64 bytes per row at a time stream at 5.2 TB/s, 128 bytes at a time at 6.1 TB/s. Row
// blocks are read without bounds checks: the packed layouts are padded to a multiple of
// 512 rows.

// Digit chunk lane l holds digit column l % 16; only the first `live` columns carry
// digits, so the other lanes skip their copy: their LDS bytes keep whatever the slot
// held, which only reaches output columns that no epilogue reads.
template <int W, int RT, int SS>
__device__ inline void thin_issue(const uint8_t* __restrict__ arow, int64_t pitch,
                                  const uint8_t* __restrict__ bstage, uint8_t* __restrict__ slot,
                                  int wave, int lane, int live) {
    using C = ThinCfg<W, RT, SS>;
#pragma unroll
    for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int h = 0; h < SS; ++h) {
            const int f = (wave * RT + q) * SS + h;
            __builtin_amdgcn_global_load_lds((const void*)(arow + plan::thin_frag_off(q, h, pitch)),
                                             (lds_void*)(slot + f * 1024), 16, 0, 0);
        }
    const bool dig = (lane & 15) < live;  // live >= 1: every 16-lane group has an active lane
    if constexpr (C::kDigKB >= W) {
#pragma unroll
        for (int d = 0; d < C::kDigPerWave; ++d) {
            const int ch = wave * C::kDigPerWave + d;
            if (dig)
                __builtin_amdgcn_global_load_lds((const void*)(bstage + plan::dig_lane(ch, lane)),
                                                 (lds_void*)(slot + C::kA + ch * 1024), 16, 0, 0);
        }
    } else {
        if (wave < C::kDigKB && dig)
            __builtin_amdgcn_global_load_lds((const void*)(bstage + plan::dig_lane(wave, lane)),
                                             (lds_void*)(slot + C::kA + wave * 1024), 16, 0, 0);
    }
}

// thin_issue for gathered rows: arq[q] is this lane's row of fragment q at super-step 0
// of the row block (the rows of a row block are any rows of the layout, listed by the
// caller), koff the stage's packed-byte offset within a row.
template <int W, int RT, int SS>
__device__ inline void thin_issue_ga(const uint8_t* const (&arq)[RT], int64_t koff,
                                     const uint8_t* __restrict__ bstage, uint8_t* __restrict__ slot,
                                     int wave, int lane, int live) {
    using C = ThinCfg<W, RT, SS>;
#pragma unroll
    for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int h = 0; h < SS; ++h) {
            const int f = (wave * RT + q) * SS + h;
            __builtin_amdgcn_global_load_lds((const void*)(arq[q] + koff + 64 * h), (lds_void*)(slot + f * 1024),
                                             16, 0, 0);
        }
    const bool dig = (lane & 15) < live;
    if constexpr (C::kDigKB >= W) {
#pragma unroll
        for (int d = 0; d < C::kDigPerWave; ++d) {
            const int ch = wave * C::kDigPerWave + d;
            if (dig)
                __builtin_amdgcn_global_load_lds((const void*)(bstage + plan::dig_lane(ch, lane)),
                                                 (lds_void*)(slot + C::kA + ch * 1024), 16, 0, 0);
        }
    } else {
        if (wave < C::kDigKB && dig)
            __builtin_amdgcn_global_load_lds((const void*)(bstage + plan::dig_lane(wave, lane)),
                                             (lds_void*)(slot + C::kA + wave * 1024), 16, 0, 0);
    }
}

template <int W, int RT, int SS>
__device__ inline void thin_mma(const uint8_t* __restrict__ slot, int wave, int lane,
                                v4i (&acc)[RT]) {
    using C = ThinCfg<W, RT, SS>;
#pragma unroll
    for (int h = 0; h < SS; ++h) {
        uint4 av[RT];
#pragma unroll
        for (int q = 0; q < RT; ++q)
            av[q] = *reinterpret_cast<const uint4*>(slot + ((wave * RT + q) * SS + h) * 1024 + lane * 16);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const v4i bv = *reinterpret_cast<const v4i*>(slot + C::kA + (4 * h + s) * 1024 + lane * 16);
#pragma unroll
            for (int q = 0; q < RT; ++q)
                acc[q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(decode16(dword_of(av[q], s)), bv, acc[q], 0, 0, 0);
        }
    }
}

// Wait until this wave's copies of the stage about to be used have landed (all but the
// copies of the NS - 2 later stages in flight), then barrier: after it, every wave's
// copies of the stage are visible in LDS. Waves that copy no digits issue fewer copies
// per stage, so their count is smaller.
// The compiler treats a copy into LDS as an event of the LDS counter as well as of the
// memory counter (with one in flight it waits lgkmcnt(0) before any LDS read), so the
// per-stage wait leaves the later stages' copies in flight on both counters (the LDS one
// capped at its 4-bit field; waiting for fewer is safe). No LDS read is outstanding
// here -- every read of the previous stage fed an MFMA -- so this is exact whether or not
// the hardware counts the copies there, and lgkmcnt(0) could drain the later copies.
constexpr int lgkm_cap(int n) { return n < 15 ? n : 15; }

template <int W, int RT, int NS, int SS>
__device__ inline void thin_stage_barrier(int wave) {
    using C = ThinCfg<W, RT, SS>;
    constexpr int kAhead = (NS - 2) * C::kOps;
    constexpr int kAheadRowOnly = (NS - 2) * C::kOpsRowOnly;
    static_assert(kAhead < 64, "vmcnt field");
    if (C::kDigKB >= W || wave < C::kDigKB)
        asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(%1)\n\ts_barrier" ::"n"(kAhead), "n"(lgkm_cap(kAhead)) : "memory");
    else
        asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(%1)\n\ts_barrier" ::"n"(kAheadRowOnly),
                     "n"(lgkm_cap(kAheadRowOnly)) : "memory");
}

// Stream-K over the units (row block, stage) of plan::sk_block_of: every block streams
// the same number of bytes, and the partial slots of each row block are exactly P, so the
// epilogue's sum does not depend on the grid size.
// GA: the rows of row block r are rows ridx[r kRB ..] of the layout (ridx lists kRB R of
// them), not rows r kRB ..; row i of the output is the product of row ridx[i].
template <int W, int RT, int NS, int SS, bool GA>
__global__ __launch_bounds__(W * 64) void k_geno_thin(const uint8_t* __restrict__ rows, int64_t pitch,
                                                      int64_t nrows, int64_t n_st, int64_t U, int64_t G,
                                                      int P, const uint8_t* __restrict__ Bq,
                                                      int32_t* __restrict__ out, int live,
                                                      const int* skip, const int32_t* __restrict__ ridx) {
    if (skip && *skip) return;
    using C = ThinCfg<W, RT, SS>;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[NS * C::kSlot];
    const int64_t b = blockIdx.x;
    const int64_t lo = plan::sk_lo(b, U, G), hi = plan::sk_lo(b + 1, U, G);
    const int64_t nt = hi - lo;
    if (nt <= 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, rr = lane & 15;
    constexpr int64_t kRB = C::kRB;  // rows per row block
    const int64_t lane_off = plan::thin_lane_off(wave, RT, rr, g, pitch);

    plan::SkCursor cur(lo, nt, n_st);
    // gathered rows: the lane's row of each fragment, looked up when the copies move to
    // another row block (once per n_st stages)
    const uint8_t* arq[RT];
    int64_t arq_r = -1;
    auto issue_next = [&](int slot) {
        if constexpr (GA) {
            if (cur.r != arq_r) {
                arq_r = cur.r;
#pragma unroll
                for (int q = 0; q < RT; ++q)
                    arq[q] = rows + (int64_t)ridx[cur.r * kRB + wave * 16 * RT + q * 16 + rr] * pitch + 16 * g;
            }
            thin_issue_ga<W, RT, SS>(arq, 64 * SS * cur.t, Bq + plan::dig_stage(cur.t, C::kB), lds + slot * C::kSlot,
                                     wave, lane, live);
        } else {
            thin_issue<W, RT, SS>(rows + plan::thin_unit_off(cur.r, kRB, cur.t, SS, pitch) + lane_off, pitch,
                                  Bq + plan::dig_stage(cur.t, C::kB), lds + slot * C::kSlot, wave, lane, live);
        }
        cur.advance();
    };
    // consume cursor
    int64_t cr = lo / n_st, ct = lo % n_st;

    v4i acc[RT];
#pragma unroll
    for (int a = 0; a < RT; ++a) acc[a] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < NS - 1; ++u) issue_next(u);
    int cs = 0;  // ring slot being consumed
    for (int64_t i = 0; i < nt; ++i) {
        thin_stage_barrier<W, RT, NS, SS>(wave);
        issue_next(cs == 0 ? NS - 1 : cs - 1);
        __builtin_amdgcn_sched_barrier(0);
        thin_mma<W, RT, SS>(lds + cs * C::kSlot, wave, lane, acc);
        cs = cs + 1 == NS ? 0 : cs + 1;
        const bool row_end = ++ct == n_st;
        if (row_end || i == nt - 1) {
            // this block's piece of row block cr is complete
            const int64_t slot = plan::sk_slot(b, cr, n_st, U, G);
#pragma unroll
            for (int a = 0; a < RT; ++a)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t r = plan::thin_out_row(cr, kRB, wave, RT, a, v, g);
                    if (r < nrows) {
                        out[plan::part_idx(slot, nrows, r, 16, rr)] = acc[a][v];
                        if (row_end)
                            for (int64_t z = slot + 1; z < P; ++z) out[plan::part_idx(z, nrows, r, 16, rr)] = 0;
                    }
                }
#pragma unroll
            for (int a = 0; a < RT; ++a) acc[a] = v4i{0, 0, 0, 0};
            // the stores count in vmcnt: let them (and the copies in flight) complete so
            // that the per-stage wait counts stay exact
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            if (row_end) {
                ct = 0;
                ++cr;
            }
        }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// ---------------------------------------------------------------- gathered pass 1
// Pass 1 (u = W^T S0 over SNPs) when S0 is zero outside a list of individuals ridx[0 ..
// ng): instead of every individual's column of the SNP-major layout, read only the listed
// individuals' rows of the individual-major layout. The contraction then runs over
// individuals, which are rows of that layout, so the MFMA operand (16 individuals of one
// SNP per lane) is a column of a staged tile: gfx950's transposing LDS read
// ds_read_b64_tr_b8 supplies it. Lane s of a 16-lane group gives the address of the 8
// bytes at (row s >> 1, half s & 1) of an 8-row x 16-byte block, and lane m receives byte
// m of those 8 rows (probed on gfx950); two reads give 16 individuals. Each byte holds 4
// SNPs, so one pair of reads feeds 4 MFMAs (SNP 4 c + j of byte c, j = 0..3).
//
// A workgroup of 8 waves owns RB bytes of the rows (4 RB SNPs; wave w the 16-byte columns
// RB / 128 w ..) and split sp of the listed individuals (64-individual stages lo .. hi),
// streamed one stage at a time: each stage copies 64 rows x RB bytes and the 16 x 64
// digit block global -> LDS into a ring of kTgNS slots, one wait + barrier per stage. The
// copy lanes take their rows' indices 16 stages at a time into registers (two groups
// held), so no index load waits inside the ring. 16-byte piece p of row r sits at slot p ^
// tg_swz(r), and digit chunk g of column n at g ^ ((n >> 2) & 3), so the transposing reads
// and the digit reads touch each LDS bank once per 16 lanes.
// Dt: digits [16 columns][ldd], column n of listed individual t at n ldd + t (0 past J).
// out[sp][s][16]: the split's int32 sums for SNP s (every split writes all its rows).
constexpr int kTgNS = 3;      // ring slots
constexpr int kTgGroup = 16;  // stages per group of row indices

typedef int v2i_t __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const uint8_t lds_cu8;

// LDS byte address of a pointer into a __shared__ array
__device__ inline uint32_t lds_addr(const uint8_t* p) { return (uint32_t)(uintptr_t)(lds_cu8*)p; }

template <int RB>
__global__ __launch_bounds__(512) void k_geno_tg(const uint8_t* __restrict__ ind, int64_t pitch, int64_t m,
                                                 const int32_t* __restrict__ ridx, int64_t ng,
                                                 const int8_t* __restrict__ Dt, int64_t ldd, int nsplit,
                                                 int32_t* __restrict__ out, const int* skip) {
    if (skip && *skip) return;
    constexpr int kTile = 64 * RB;            // staged rows
    constexpr int kSlot = kTile + 1024;       // + digits
    constexpr int kCopies = plan::tg_copies<RB>();  // row copies per wave and stage (= its columns)
    static_assert(RB == 128 || RB == 256, "128- or 256-byte rows");
    __shared__ __attribute__((aligned(1024))) uint8_t lds[kTgNS * kSlot];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t b0 = (int64_t)blockIdx.x * RB;  // byte offset of the workgroup's SNPs
    const int sp = blockIdx.y;
    const int64_t nst = ng / 64;
    const int64_t lo = plan::tg_split_lo(sp, nst, nsplit), hi = plan::tg_split_lo(sp + 1, nst, nsplit);
    const int64_t nt = hi - lo;
    // copy roles: copy instruction i of wave w fills LDS bytes (8 i + w) KB: stage row
    // cr[i], piece slot lane % pieces (the global piece swizzled into it, clamped inside
    // the row for a last workgroup past the pitch); lanes 0..7 copy a chunk of a digit
    // column
    int cr[kCopies];
    const uint8_t* ibase[kCopies];
#pragma unroll
    for (int i = 0; i < kCopies; ++i) {
        cr[i] = plan::tg_copy_row<RB>(i, w, lane);
        ibase[i] = ind + plan::tg_piece_off<RB>(b0, cr[i], lane, pitch);
    }
    const int8_t* const dbase = Dt + (lane < 8 ? plan::tg_dig_off(w, lane, ldd) : 0);
    v4i acc[kCopies][4];
#pragma unroll
    for (int i = 0; i < kCopies; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4i{0, 0, 0, 0};
    const int g = lane >> 4, sl = lane & 15;
    if (nt > 0) {
        // the copy rows' indices for stage t of the split (the last stage's past the end)
        auto row_of = [&](int64_t t, int i) -> int32_t {
            return ridx[(lo + (t < nt ? t : nt - 1)) * 64 + cr[i]];
        };
        auto issue = [&](int64_t t, const int32_t (&r)[kCopies], int slot) {
            const int64_t t0 = (lo + (t < nt ? t : nt - 1)) * 64;
            uint8_t* dst = lds + slot * kSlot;
#pragma unroll
            for (int i = 0; i < kCopies; ++i)
                __builtin_amdgcn_global_load_lds((const void*)(ibase[i] + (int64_t)r[i] * pitch),
                                                 (lds_void*)(dst + 1024 * (8 * i + w)), 16, 0, 0);
            if (lane < 8)
                __builtin_amdgcn_global_load_lds((const void*)(dbase + t0), (lds_void*)(dst + kTile + 128 * w),
                                                 16, 0, 0);
        };
        int32_t cur[kTgGroup][kCopies], nxt[kTgGroup][kCopies];
#pragma unroll
        for (int q = 0; q < kTgGroup; ++q)
#pragma unroll
            for (int i = 0; i < kCopies; ++i) {
                cur[q][i] = row_of(q, i);
                nxt[q][i] = row_of(kTgGroup + q, i);
            }
#pragma unroll
        for (int u = 0; u < kTgNS - 1; ++u) issue(u, cur[u], u);
        int cs = 0;
        for (int64_t t0 = 0; t0 < nt; t0 += kTgGroup) {
#pragma unroll
            for (int q = 0; q < kTgGroup; ++q) {
                const int64_t t = t0 + q;
                if (t >= nt) continue;  // uniform in the workgroup
                // this wave's copies of stage t have landed (the later stage's are in
                // flight); after the barrier every wave's have, and no wave still reads the
                // slot the next copies overwrite
                asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"((kCopies + 1) * (kTgNS - 2)) : "memory");
                constexpr int ahead = kTgNS - 1;
                if (q + ahead < kTgGroup)
                    issue(t + ahead, cur[(q + ahead) % kTgGroup], cs == 0 ? kTgNS - 1 : cs - 1);
                else
                    issue(t + ahead, nxt[(q + ahead) % kTgGroup], cs == 0 ? kTgNS - 1 : cs - 1);
                __builtin_amdgcn_sched_barrier(0);
                const uint8_t* slot = lds + cs * kSlot;
                const v4i bv = *reinterpret_cast<const v4i*>(slot + kTile + plan::tg_dig_slot(sl, g));
                const int rr = 16 * g + (sl >> 1);
#pragma unroll
                for (int i = 0; i < kCopies; ++i) {
                    const int pc = kCopies * w + i;  // the 16-byte column of every row
                    // written as asm: the compiler would otherwise wait for every copy in
                    // flight (vmcnt(0)) before a transposing read, taking it for a read of
                    // the copies' destination; both reads complete inside the block
                    const uint32_t a1 = lds_addr(slot + plan::tg_row_slot<RB>(rr, pc) + 8 * (sl & 1));
                    const uint32_t a2 = lds_addr(slot + plan::tg_row_slot<RB>(rr + 8, pc) + 8 * (sl & 1));
                    v2i_t lo8, hi8;
                    asm volatile("ds_read_b64_tr_b8 %0, %2\n\tds_read_b64_tr_b8 %1, %3\n\ts_waitcnt lgkmcnt(0)"
                                 : "=&v"(lo8), "=&v"(hi8)
                                 : "v"(a1), "v"(a2)
                                 : "memory");
                    const uint32_t wd[4] = {(uint32_t)lo8[0], (uint32_t)lo8[1], (uint32_t)hi8[0], (uint32_t)hi8[1]};
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        v4i av;
#pragma unroll
                        for (int e = 0; e < 4; ++e) av[e] = (int)((wd[e] >> (2 * j)) & kFieldMask);
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc[i][j], 0, 0, 0);
                    }
                }
                cs = cs + 1 == kTgNS ? 0 : cs + 1;
            }
            // the next group's indices move up; the one after is read now, 16 stages
            // before its first use
#pragma unroll
            for (int q = 0; q < kTgGroup; ++q)
#pragma unroll
                for (int i = 0; i < kCopies; ++i) {
                    cur[q][i] = nxt[q][i];
                    nxt[q][i] = row_of(t0 + 2 * kTgGroup + q, i);
                }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    // accumulator element v of lane (g, column sl) for column piece pc: MFMA row 4 g + v =
    // byte 16 pc + 4 g + v of the workgroup's RB, SNP 4 (byte) + j
#pragma unroll
    for (int i = 0; i < kCopies; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t snp = plan::tg_out_snp<RB>(b0, w, i, g, v, j);
                if (snp < m) out[plan::part_idx(sp, m, snp, 16, sl)] = acc[i][j][v];
            }
}

// ---------------------------------------------------------------- wide kernel
// Block shape <W waves, NS ring stages, CG column groups, RA row fragments per wave>:
// the waves form W / CG row groups x CG column groups; a wave owns RA row fragments of
// 32 rows and CB = JT / CG column fragments of 32, so a block covers
// (W / CG) 32 RA rows x all 32 JT digit columns. Per stage a wave reads RA KB of
// packed rows (4 K-steps each) and 4 CB KB of digits from LDS for 4 RA CB MFMAs, so
// the taller tiles (RA = 4) need a third less LDS traffic per MFMA than RA = 2.

// Issue the global -> LDS copies of one stage into ring slot `slot`: kAPerWave row
// fragments (lane l copies the 16 packed bytes of row l % 32, half l / 32; they land
// lane-ordered) and kBPerWave digit chunks. With fewer chunks than waves the spare
// waves copy into a private sink, so every wave issues the same number of copies and
// one vmcnt fits all. arow: this lane's row pointer of fragment 0 at the stage.
template <class C>
__device__ inline void wide_issue(const uint8_t* __restrict__ arow, int64_t frag_stride,
                                  const uint8_t* __restrict__ bstage, uint8_t* __restrict__ slot,
                                  uint8_t* __restrict__ sink, int wave, int lane) {
#pragma unroll
    for (int q = 0; q < C::kAPerWave; ++q) {
        const int f = wave + C::kW * q;
        __builtin_amdgcn_global_load_lds((const void*)(arow + f * frag_stride),
                                         (lds_void*)(slot + f * 1024), 16, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < C::kBPerWave; ++q) {
        const int chunk = wave + C::kW * q;
        if (chunk < C::kBChunks)
            __builtin_amdgcn_global_load_lds((const void*)(bstage + plan::dig_lane(chunk, lane)),
                                             (lds_void*)(slot + C::kA + chunk * 1024), 16, 0, 0);
        else
            __builtin_amdgcn_global_load_lds((const void*)(bstage + plan::dig_lane(chunk % C::kBChunks, lane)),
                                             (lds_void*)(sink + (chunk - C::kBChunks) * 1024), 16, 0, 0);
    }
}

// 4 K-steps of one stage from LDS: the wave's RA row fragments, then its digit
// fragments -- one K-step ahead of the MFMAs that use them (PF = 0), or two K-steps
// ahead with the reads pinned before the MFMAs they overlap (PF = 1).
template <class C, int PF>
__device__ inline void wide_mma(const uint8_t* __restrict__ slot, int rgroup, int cgroup, int lane,
                                v16i (&acc)[4][4]) {
    constexpr int RA = C::kRA, CB = C::kCB, JT = C::kJT;
    uint4 av[RA];
#pragma unroll
    for (int a = 0; a < RA; ++a)
        av[a] = *reinterpret_cast<const uint4*>(slot + (rgroup * RA + a) * 1024 + lane * 16);
    const uint8_t* cl = slot + C::kA + (cgroup * CB) * 1024 + lane * 16;
    if constexpr (PF == 1) {
        // digits two K-steps ahead: at most 2 + 3 CB reads in flight (the LDS counter
        // tracks 15), each issued before the MFMAs it runs behind
        v4i bf[4][CB];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int b = 0; b < CB; ++b) bf[s][b] = *reinterpret_cast<const v4i*>(cl + (s * JT + b) * 1024);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s + 2 < 4)
#pragma unroll
                for (int b = 0; b < CB; ++b)
                    bf[s + 2][b] = *reinterpret_cast<const v4i*>(cl + ((s + 2) * JT + b) * 1024);
            // keep the reads ahead of this K-step's MFMAs (the scheduler sinks them to their uses)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int a = 0; a < RA; ++a) {
                const v4i ad = decode16(dword_of(av[a], s));
#pragma unroll
                for (int b = 0; b < CB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ad, bf[s][b], acc[a][b], 0, 0, 0);
            }
        }
        return;
    }
    v4i bcur[CB];
#pragma unroll
    for (int b = 0; b < CB; ++b) bcur[b] = *reinterpret_cast<const v4i*>(cl + b * 1024);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        v4i bnx[CB];
        if (s < 3)
#pragma unroll
            for (int b = 0; b < CB; ++b)
                bnx[b] = *reinterpret_cast<const v4i*>(cl + ((s + 1) * JT + b) * 1024);
#pragma unroll
        for (int a = 0; a < RA; ++a) {
            const v4i ad = decode16(dword_of(av[a], s));
#pragma unroll
            for (int b = 0; b < CB; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ad, bcur[b], acc[a][b], 0, 0, 0);
        }
        if (s < 3)
#pragma unroll
            for (int b = 0; b < CB; ++b) bcur[b] = bnx[b];
    }
}

// 256-genotype stages (KS = 2, plan::wide2_*): each of the wave's row fragments takes two
// copy instructions of 16 rows x 64 packed bytes; abase is the block's first row at its
// first stage (uniform), aoff the lane's 32-bit offsets from it.
template <class C>
__device__ inline void wide2_issue(const uint8_t* __restrict__ abase, const uint32_t (&aoff)[C::kAPerWave][2],
                                   const uint8_t* __restrict__ bbase, uint32_t boff, uint8_t* __restrict__ slot,
                                   uint8_t* __restrict__ sink, int wave, int64_t t) {
    const uint8_t* at = abase + plan::wide2_stage_step(t);
#pragma unroll
    for (int q = 0; q < C::kAPerWave; ++q) {
        const int f = wave + C::kW * q;
#pragma unroll
        for (int i = 0; i < 2; ++i)
            __builtin_amdgcn_global_load_lds((const void*)(at + aoff[q][i]), (lds_void*)(slot + f * 2048 + i * 1024),
                                             16, 0, 0);
    }
    const uint8_t* bt = bbase + plan::dig_stage(t, C::kB) + boff;
#pragma unroll
    for (int q = 0; q < C::kBPerWave; ++q) {
        const int chunk = wave + C::kW * q;
        if (chunk < C::kBChunks)
            __builtin_amdgcn_global_load_lds((const void*)(bt + chunk * 1024), (lds_void*)(slot + C::kA + chunk * 1024),
                                             16, 0, 0);
        else
            __builtin_amdgcn_global_load_lds((const void*)(bt + (chunk % C::kBChunks) * 1024),
                                             (lds_void*)(sink + (chunk - C::kBChunks) * 1024), 16, 0, 0);
    }
}

// 8 K-steps of a 256-genotype stage: K-block kb (0, 1) of row fragment a is chunk
// 2 kb + g of row rr (the 16 packed bytes that K-steps 4 kb .. 4 kb + 3 of the 128-genotype
// layout read); digit chunk of K-step ks, column fragment b: ks JT + b. Digits two K-steps
// ahead, pinned as in wide_mma<C, 1>.
template <class C>
__device__ inline void wide2_mma(const uint8_t* __restrict__ slot, int rgroup, int cgroup, int lane,
                                 v16i (&acc)[4][4]) {
    constexpr int RA = C::kRA, CB = C::kCB, JT = C::kJT;
    const int rr = lane & 31, g = lane >> 5;
    const uint8_t* cl = slot + C::kA + (cgroup * CB) * 1024 + lane * 16;
    v4i bf[8][CB];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int b = 0; b < CB; ++b) bf[s][b] = *reinterpret_cast<const v4i*>(cl + (s * JT + b) * 1024);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        uint4 av[RA];
#pragma unroll
        for (int a = 0; a < RA; ++a)
            av[a] = *reinterpret_cast<const uint4*>(slot + (rgroup * RA + a) * 2048 + plan::wide2_lds_chunk(rr, 2 * kb + g));
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int ks = 4 * kb + s;
            if (ks + 2 < 8)
#pragma unroll
                for (int b = 0; b < CB; ++b)
                    bf[ks + 2][b] = *reinterpret_cast<const v4i*>(cl + ((ks + 2) * JT + b) * 1024);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int a = 0; a < RA; ++a) {
                const v4i ad = decode16(dword_of(av[a], s));
#pragma unroll
                for (int b = 0; b < CB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ad, bf[ks][b], acc[a][b], 0, 0, 0);
            }
        }
    }
}

// Wait until this wave's copies of the stage about to be multiplied have landed (at
// most the copies of the NS - 2 later stages still in flight) and its LDS reads of
// the previous stage are complete, then synchronize the block, so that every wave's
// copies are visible and the slot refilled next is no longer read.
template <class C, int NS>
__device__ inline void wide_stage_barrier() {
    constexpr int kAhead = (NS - 2) * C::kOps;
    static_assert(kAhead < 64, "vmcnt field");
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(%1)\n\ts_barrier" ::"n"(kAhead), "n"(lgkm_cap(kAhead)) : "memory");
}

// All operands pass through LDS (no loaded value lives in VGPRs across a stage), so
// the only waits on the copies are the explicit per-stage ones above.
template <int JT, int W, int NS, int CG, int RA, int PF, int KS>
__global__ __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(2))) void k_geno_wide(
    const uint8_t* __restrict__ rows, int64_t pitch, int64_t nrows, int64_t n_ss, int64_t ss_per_split,
    int64_t R, int S, const uint8_t* __restrict__ Bq, int32_t* __restrict__ out, const int* skip) {
    if (skip && *skip) return;
    using C = WideCfg<JT, W, CG, RA, KS>;
    constexpr int CB = C::kCB;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[NS * C::kSlot + C::kSink];
    const plan::WideBlock wb = plan::wide_block(blockIdx.x, R, S, C::kRows, n_ss, ss_per_split);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rgroup = wave / CG, cgroup = wave % CG;
    const int g = lane >> 5, rr = lane & 31;
    const int64_t blk_row = wb.blk_row, H0 = wb.H0, nt = wb.nt / KS;  // stages of 128 KS genotypes
    // KS = 1: fragment f of the block = rows blk_row + 32 f .. + 31, lane row pointer
    const uint8_t* arow = rows + plan::wide_row_base(blk_row, rr, g, H0, pitch);
    const int64_t fstride = plan::wide_frag_stride(pitch);
    // KS = 2: uniform block base and the lane's 32-bit offsets (the launch checks that a
    // block's rows span less than 4 GB)
    const uint8_t* abase = rows + plan::wide_row_base(blk_row, 0, 0, H0, pitch);
    uint32_t aoff[C::kAPerWave][2];
    if constexpr (KS == 2) {
#pragma unroll
        for (int q = 0; q < C::kAPerWave; ++q)
#pragma unroll
            for (int i = 0; i < 2; ++i) aoff[q][i] = (uint32_t)plan::wide2_lane_off(wave + W * q, i, lane, pitch);
    }
    const uint8_t* bst = Bq + plan::dig_stage(H0 / KS, C::kB);
    const uint32_t boff = (uint32_t)(lane * 16);
    uint8_t* sink = lds + NS * C::kSlot;
    auto issue = [&](int64_t t, uint8_t* slot) {
        if constexpr (KS == 1)
            wide_issue<C>(arow + plan::wide_stage_step(t), fstride, bst + plan::dig_stage(t, C::kB), slot, sink, wave,
                          lane);
        else
            wide2_issue<C>(abase, aoff, bst, boff, slot, sink, wave, t);
    };

    v16i acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0;
    if (nt > 0) {
        // prologue: stages 0 .. NS - 2 (clamped to the last stage; a clamped copy
        // lands in a slot that is never multiplied)
#pragma unroll
        for (int u = 0; u < NS - 1; ++u) issue(u < nt ? u : nt - 1, lds + u * C::kSlot);
        for (int64_t t0 = 0; t0 < nt; t0 += NS) {
#pragma unroll
            for (int u = 0; u < NS; ++u) {
                const int64_t t = t0 + u;
                if (t >= nt) break;
                wide_stage_barrier<C, NS>();
                // refill the slot multiplied in the previous stage with stage t + NS - 1,
                // issued before this stage's MFMAs (the scheduler may not sink it)
                issue(t + NS - 1 < nt ? t + NS - 1 : nt - 1, lds + ((u + NS - 1) % NS) * C::kSlot);
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (KS == 1)
                    wide_mma<C, PF>(lds + u * C::kSlot, rgroup, cgroup, lane, acc);
                else
                    wide2_mma<C>(lds + u * C::kSlot, rgroup, cgroup, lane, acc);
            }
        }
        // no copy into LDS may be in flight when the block ends
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    constexpr int J16 = 32 * JT;
    int32_t* o = out + plan::part_idx(wb.sp, nrows, 0, J16, 0);
    const int64_t row_base = blk_row + rgroup * 32 * RA;
#pragma unroll
    for (int a = 0; a < RA; ++a)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t r = plan::wide_out_row(row_base, a, v, g);
            if (r < nrows)
#pragma unroll
                for (int b = 0; b < CB; ++b) o[plan::part_idx(0, nrows, r, J16, (cgroup * CB + b) * 32 + rr)] = acc[a][b][v];
        }
}

// ---------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}
__device__ inline double block_sum(double v) {
    __shared__ double sh[4];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
    return r;
}
__device__ inline double block_max(double v) {
    __shared__ double sh[4];
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    return r;
}

// per-column max |v| partials
__global__ __launch_bounds__(256) void k_colmax_parts(int64_t rows, const double* __restrict__ v,
                                                      int64_t ldv, double* part, const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    double m = 0.0;
    for (int64_t q = blockIdx.x * 256ll + threadIdx.x; q < rows; q += (int64_t)gridDim.x * 256)
        m = fmax(m, fabs(v[c * ldv + q]));
    const double r = block_max(m);
    if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = r;
}

// One wave per column: max (want_max) or ordered sum of the column's partials; the
// scale exponent sh_c = (8 L_c - 2) - e_c with max|v| < 2^e_c when sh != nullptr.
// One wave per column (16 waves): each lane loads its up to 16 partials at once (so the
// loads are in flight together) and adds them in the order j = lane, lane + 64, ..., then
// the wave reduces; slots j >= nparts read as 0, which changes neither the maximum of
// non-negative values nor the sum.
constexpr int kFinishThreads = 1024;
__global__ __launch_bounds__(kFinishThreads) void k_finish_cols(const double* __restrict__ part, int nparts,
                                                                int k, int want_max, const LimbMap lm, int* sh,
                                                                double* out, const int* skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    constexpr int kPer = kMaxParts / 64;
    for (int c = threadIdx.x >> 6; c < k; c += kFinishThreads / 64) {
        double x[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int j = lane + 64 * i;
            x[i] = j < nparts ? part[c * kMaxParts + j] : 0.0;
        }
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) v = want_max ? fmax(v, x[i]) : v + x[i];
        v = want_max ? wave_max(v) : wave_sum(v);
        if (lane == 0) {
            if (out) out[c] = v;
            if (sh) {
                int e = 0;
                if (v > 0.0) (void)frexp(v, &e);  // v < 2^e
                sh[c] = (v > 0.0 && isfinite(v)) ? (8 * lm.L[c] - 2) - e : 0;
            }
        }
    }
}
// Digit matrix of k columns: one thread per genotype kap of column c = blockIdx.y (the
// reads of v are coalesced). It rounds the value to the column's block fixed point once
// (q = round(z 2^sh_c)), splits q into its L_c balanced base-256 digits and stores digit d
// at plan::bq_index of digit column off_c + d; the 16 genotypes of a 16-byte chunk are
// 16 consecutive threads, so each store instruction writes whole chunks. Genotypes
// K .. Kpad - 1 get digit 0; the last grid row (blockIdx.y == k) zeroes the unused digit
// columns J .. J16 - 1. part[c][block]: the block's sum of wts[kap] q 2^-sh_c (wts = 1
// when null), in a fixed order.
__device__ inline int balanced_digit(long long& t) {
    int b = (int)(t & 0xFF);
    if (b >= 128) b -= 256;
    t = (t - b) >> 8;
    return b;
}
__global__ __launch_bounds__(256) void k_quantize(const double* __restrict__ v, int64_t ldv,
                                                  int64_t K, int64_t Kpad, int k, const LimbMap lm,
                                                  int nfrag, int wide, int J, int J16,
                                                  const int* __restrict__ sh,
                                                  const double* __restrict__ wts,
                                                  uint8_t* __restrict__ Bq, double* part,
                                                  const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const int fw = wide ? 32 : 16;
    double acc = 0.0;
    for (int64_t kap = blockIdx.x * 256ll + threadIdx.x; kap < Kpad; kap += (int64_t)gridDim.x * 256) {
        if (c == k) {
            for (int j = J; j < J16; ++j) Bq[plan::bq_index(kap, j % fw, j / fw, nfrag, wide != 0)] = 0;
            continue;
        }
        const int L = lm.L[c], off = lm.off[c], e = sh[c];
        long long q = 0;
        if (kap < K) {
            q = llrint(ldexp(v[c * ldv + kap], e));
            acc += (wts ? wts[kap] : 1.0) * ldexp((double)q, -e);
        }
        for (int d = 0; d < L; ++d) {
            const int j = off + d;
            Bq[plan::bq_index(kap, j % fw, j / fw, nfrag, wide != 0)] = (uint8_t)(int8_t)balanced_digit(q);
        }
    }
    if (c < k) {
        const double r = block_sum(acc);
        if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = r;
    }
}
// Dt[n][t] (n < 16, t < ng): digit lm.dig[n] of column lm.col[n] of round(S0[c][ridx[t]]
// 2^sh_c), the digits k_quantize forms for that individual; 0 for n >= J.
__global__ __launch_bounds__(256) void k_quantize_gather(const double* __restrict__ v, int64_t ldv,
                                                         const int32_t* __restrict__ ridx, int64_t ng,
                                                         const LimbMap lm, int J, const int* __restrict__ sh,
                                                         int8_t* __restrict__ Dt, int64_t ldd, const int* skip) {
    if (skip && *skip) return;
    const int nn = blockIdx.y;
    for (int64_t t = blockIdx.x * 256ll + threadIdx.x; t < ng; t += (int64_t)gridDim.x * 256) {
        int dig = 0;
        if (nn < J) {
            const int c = lm.col[nn], d = lm.dig[nn];
            long long q = llrint(ldexp(v[c * ldv + ridx[t]], sh[c]));
            for (int a = 0; a <= d; ++a) dig = balanced_digit(q);
        }
        Dt[nn * ldd + t] = (int8_t)dig;
    }
}

// blocks per column of k_quantize
inline int quantize_parts(int64_t Kpad) { return parts_for(Kpad); }
__global__ __launch_bounds__(256) void k_tune_fill(double* __restrict__ v, int64_t count, uint64_t seed) {
    for (int64_t q = blockIdx.x * 256ll + threadIdx.x; q < count; q += (int64_t)gridDim.x * 256) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(q + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        v[q] = (double)(int64_t)(z >> 11) * 0x1.0p-52 - 1.0;
    }
}

// Pass 1 of G e_j for the unit vectors e_j0 .. e_j0+k-1 without a contraction: W^T e_j
// is individual j's standardised row, so u[c][s] = inv_s^2 / M (g_{s, j0+c} - 2 f_s), the
// pass-1 epilogue's value (sig2g = 1, T = 1). g from the individual-major dosage rows.
__global__ __launch_bounds__(256) void k_unit_u(const uint8_t* __restrict__ ind, int64_t pitch, int64_t j0,
                                                int64_t m, const double* __restrict__ twof,
                                                const double* __restrict__ inv, double inv_m,
                                                double* __restrict__ U) {
    const int c = blockIdx.y;
    const uint8_t* row = ind + (j0 + c) * pitch;
    for (int64_t q = blockIdx.x * 256ll + threadIdx.x; q < m; q += (int64_t)gridDim.x * 256) {
        const int g = (row[q >> 2] >> (2 * (q & 3))) & 3;
        const double iv = inv[q];
        U[c * m + q] = inv_m * iv * iv * ((double)g - twof[q]);
    }
}

// Epilogues: a block takes 64 rows x k columns. Phase 1 reads the digit sums with
// consecutive threads on consecutive digits (coalesced), phase 2 writes the FP64
// results column-major with consecutive threads on consecutive rows.
// Epilogue tiles of kEpiRows rows. The int32 partial sums of the S splits for a tile's
// rows are one contiguous run of kEpiRows J16 words per split: each thread adds 16
// consecutive words over the splits in int64 (exact, so the order does not matter) with
// four 16-byte loads per split, the sums go to LDS, and one thread per (row, column)
// combines the column's L digit sums in FP64 in digit order, v += s_l 2^(8 l).
constexpr int kEpiRows = 32;
constexpr int kEpiWords = 16;  // partial words per thread and split

__global__ __launch_bounds__(256) void k_epilogue(const int32_t* __restrict__ out, int nsplit,
                                                  int64_t rows, int J16, const LimbMap lm, int k,
                                                  const int* __restrict__ sh, int mode,
                                                  const double* __restrict__ TC,
                                                  const double* __restrict__ twof,
                                                  const double* __restrict__ inv,
                                                  const double* __restrict__ th, double inv_m,
                                                  double* __restrict__ res, int64_t ldr,
                                                  double* part, const int* skip,
                                                  const int32_t* __restrict__ ridx) {
    if (skip && *skip) return;
    __shared__ long long dsum[kEpiRows * 128];
    __shared__ double tile[kEpiRows][kMaxCols + 1];
    __shared__ double cmax[kMaxCols];
    const int64_t split_stride = rows * J16;
    const int64_t ntiles = (rows + kEpiRows - 1) / kEpiRows;
    // J16 is a multiple of 16 (16 or 32 JT), so 16 words never straddle a row
    const int nthr = kEpiRows * J16 / kEpiWords;
    for (int c = threadIdx.x; c < k; c += 256) cmax[c] = 0.0;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t r0 = t * kEpiRows;
        __syncthreads();
        if (threadIdx.x < nthr) {
            const int w0 = threadIdx.x * kEpiWords;
            long long sm[kEpiWords];
#pragma unroll
            for (int i = 0; i < kEpiWords; ++i) sm[i] = 0;
            if (r0 + w0 / J16 < rows) {
                const int32_t* src = out + r0 * J16 + w0;
                for (int sp = 0; sp < nsplit; ++sp) {
                    const int4* q = reinterpret_cast<const int4*>(src + sp * split_stride);
                    const int4 a = q[0], b = q[1], c4 = q[2], d = q[3];
                    sm[0] += a.x; sm[1] += a.y; sm[2] += a.z; sm[3] += a.w;
                    sm[4] += b.x; sm[5] += b.y; sm[6] += b.z; sm[7] += b.w;
                    sm[8] += c4.x; sm[9] += c4.y; sm[10] += c4.z; sm[11] += c4.w;
                    sm[12] += d.x; sm[13] += d.y; sm[14] += d.z; sm[15] += d.w;
                }
            }
#pragma unroll
            for (int i = 0; i < kEpiWords; ++i) dsum[w0 + i] = sm[i];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kEpiRows * k; q += 256) {
            const int rl = q / k, c = q % k;
            const int64_t r = r0 + rl;
            double val = 0.0;
            if (r < rows) {
                const long long* ds = dsum + rl * J16 + lm.off[c];
                double v = 0.0;
                for (int l = 0; l < lm.L[c]; ++l) v += ldexp((double)ds[l], 8 * l);
                const double P = ldexp(v, -sh[c]);
                if (mode == 0) {
                    const double iv = inv[r];
                    val = th[c * 5 + 1] * inv_m * iv * iv * (P - twof[r] * TC[c]);
                } else {
                    val = P - TC[c];
                }
            }
            tile[rl][c] = val;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kEpiRows * k; q += 256) {
            const int c = q / kEpiRows, rl = q % kEpiRows;
            const int64_t r = r0 + rl;
            if (r < rows) res[c * ldr + (ridx ? (int64_t)ridx[r] : r)] = tile[rl][c];
        }
        if (mode == 0) {
            // running per-column maxima (one wave per column, fixed order)
            for (int c = threadIdx.x >> 6; c < k; c += 4) {
                double m = 0.0;
                const int rl = threadIdx.x & 63;
                if (rl < kEpiRows && r0 + rl < rows) m = fabs(tile[rl][c]);
                m = wave_max(m);
                if ((threadIdx.x & 63) == 0) cmax[c] = fmax(cmax[c], m);
            }
        }
    }
    __syncthreads();
    if (mode == 0 && part)
        for (int c = threadIdx.x; c < k; c += 256) part[c * kMaxParts + blockIdx.x] = cmax[c];
}

using plan::kThin;
using plan::kWide;
using plan::rows_per_block;
using plan::ThinShape;
using plan::Variant;
using plan::WideShape;
using plan::wide_fits;

template <int Q>
void launch_thin_q(hipStream_t s, const Variant& v, const uint8_t* rows, int64_t pitch, int64_t nrows,
                   int64_t nss, int64_t R, const uint8_t* Bq, int32_t* out, int live, const int* skip,
                   const int32_t* ridx) {
    constexpr ThinShape t = kThin[Q];
    const int64_t nst = nss / t.SS;
    if (ridx)
        k_geno_thin<t.W, t.RT, t.NS, t.SS, true><<<(unsigned)v.G, t.W * 64, 0, s>>>(
            rows, pitch, nrows, nst, R * nst, v.G, v.S, Bq, out, live, skip, ridx);
    else
        k_geno_thin<t.W, t.RT, t.NS, t.SS, false><<<(unsigned)v.G, t.W * 64, 0, s>>>(
            rows, pitch, nrows, nst, R * nst, v.G, v.S, Bq, out, live, skip, nullptr);
}

template <int Q>
int thin_blocks_per_cu() {
    constexpr ThinShape t = kThin[Q];
    int nb = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_geno_thin<t.W, t.RT, t.NS, t.SS, false>,
                                                           t.W * 64, 0));
    return nb;
}

// resident blocks per CU of thin shape cfg (a property of the compiled kernel: queried once)
int thin_occupancy(int cfg) {
    static int cache[plan::kNumThin] = {0};
    if (cache[cfg] > 0) return cache[cfg];
    int nb = 0;
    switch (cfg) {
        case 0: nb = thin_blocks_per_cu<0>(); break;
        case 1: nb = thin_blocks_per_cu<1>(); break;
        case 2: nb = thin_blocks_per_cu<2>(); break;
        case 3: nb = thin_blocks_per_cu<3>(); break;
        case 4: nb = thin_blocks_per_cu<4>(); break;
        case 5: nb = thin_blocks_per_cu<5>(); break;
        case 6: nb = thin_blocks_per_cu<6>(); break;
        default: nb = thin_blocks_per_cu<7>(); break;
    }
    cache[cfg] = std::max(nb, 1);
    return cache[cfg];
}

template <int JT, int Q>
void launch_wide_q(hipStream_t s, unsigned grid, const uint8_t* rows, int64_t pitch, int64_t nrows,
                   int64_t nss, int64_t per, int64_t R, int S, const uint8_t* Bq, int32_t* out,
                   const int* skip) {
    if constexpr (wide_fits(JT, Q)) {
        constexpr WideShape w = kWide[Q];
        // 256-genotype stages address a block's rows with 32-bit lane offsets
        require(w.KS == 1 || (int64_t)rows_per_block(true, Q) * pitch < ((int64_t)1 << 32), GLOBE_E_INVALID,
                "rows of a contraction block span 4 GB");
        k_geno_wide<JT, w.W, w.NS, w.CG, w.RA, w.PF, w.KS><<<grid, w.W * 64, 0, s>>>(rows, pitch, nrows, nss, per,
                                                                                    R, S, Bq, out, skip);
    } else {
        require(false, GLOBE_E_INVALID, "wide contraction shape not available for this width");
    }
}

template <int JT>
void launch_wide(hipStream_t s, const Variant& v, unsigned grid, const uint8_t* rows, int64_t pitch,
                 int64_t nrows, int64_t nss, int64_t per, int64_t R, const uint8_t* Bq, int32_t* out,
                 const int* skip) {
    switch (v.cfg) {
        case 0: launch_wide_q<JT, 0>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 1: launch_wide_q<JT, 1>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 2: launch_wide_q<JT, 2>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 3: launch_wide_q<JT, 3>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 4: launch_wide_q<JT, 4>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 5: launch_wide_q<JT, 5>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 6: launch_wide_q<JT, 6>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 7: launch_wide_q<JT, 7>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 8: launch_wide_q<JT, 8>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 9: launch_wide_q<JT, 9>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 10: launch_wide_q<JT, 10>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 11: launch_wide_q<JT, 11>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        case 12: launch_wide_q<JT, 12>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
        default: launch_wide_q<JT, 13>(s, grid, rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip); break;
    }
}

void launch_variant(hipStream_t s, bool wide, int JT, const Variant& v, const uint8_t* rows,
                    int64_t pitch, int64_t nrows, int64_t nss, const uint8_t* Bq, int32_t* out, int live,
                    const int* skip, const int32_t* ridx = nullptr) {
    const int64_t R = (nrows + rows_per_block(wide, v.cfg) - 1) / rows_per_block(wide, v.cfg);
    const int64_t per = (nss + v.S - 1) / v.S;
    const unsigned grid = (unsigned)(R * v.S);
    require(!ridx || !wide, GLOBE_E_INVALID, "gathered rows take the thin contraction only");
    if (!wide) {
        require(nss % kThin[v.cfg].SS == 0 && v.G >= 1 && v.G <= R * (nss / kThin[v.cfg].SS),
                GLOBE_E_INVALID, "thin shape needs whole stages per row and at most one block per unit");
        switch (v.cfg) {
            case 0: launch_thin_q<0>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
            case 1: launch_thin_q<1>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
            case 2: launch_thin_q<2>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
            case 3: launch_thin_q<3>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
            case 4: launch_thin_q<4>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
            case 5: launch_thin_q<5>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
            case 6: launch_thin_q<6>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
            default: launch_thin_q<7>(s, v, rows, pitch, nrows, nss, R, Bq, out, live, skip, ridx); break;
        }
    } else if (JT == 1) {
        launch_wide<1>(s, v, grid, rows, pitch, nrows, nss, per, R, Bq, out, skip);
    } else if (JT == 2) {
        launch_wide<2>(s, v, grid, rows, pitch, nrows, nss, per, R, Bq, out, skip);
    } else if (JT == 3) {
        launch_wide<3>(s, v, grid, rows, pitch, nrows, nss, per, R, Bq, out, skip);
    } else {
        launch_wide<4>(s, v, grid, rows, pitch, nrows, nss, per, R, Bq, out, skip);
    }
    HIP_CHECK(hipGetLastError());
}

}  // namespace

// Tile autotuner: the first contraction of each shape (pass, rows, K, digit columns)
// times every candidate variant of plan::candidates once on the live data with HIP
// events and keeps the fastest; every variant computes the same exact integer sums, so
// the choice never changes a result.
// The candidate lists and a product's geometry are pure functions of the shape (and of the
// compiled kernels' occupancy, fixed for the process). The thin lists walk every row block
// for each grid, about 0.7 ms of host time per product at N = 200k, so both are computed
// once per shape and kept for the process.
static std::mutex g_plan_mu;
static std::vector<Variant> candidates(bool wide, int JT, int64_t nrows, int64_t nss, int ncu) {
    static std::map<std::tuple<bool, int, int64_t, int64_t, int>, std::vector<Variant>> memo;
    const auto key = std::make_tuple(wide, wide ? JT : 0, nrows, nss, ncu);
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
    }
    int bpc[plan::kNumThin];
    for (int q = 0; q < plan::kNumThin; ++q) bpc[q] = wide ? 1 : thin_occupancy(q);
    std::vector<Variant> v = plan::candidates(wide, JT, nrows, nss, ncu, bpc);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    return memo.emplace(key, std::move(v)).first->second;
}
static plan::Geometry product_geometry(int J, int64_t m, int64_t n, int64_t pitch_s, int64_t pitch_i, int64_t r0,
                                       int ncu) {
    static std::map<std::tuple<int, int64_t, int64_t, int64_t, int64_t, int64_t, int>, plan::Geometry> memo;
    const auto key = std::make_tuple(J, m, n, pitch_s, pitch_i, r0, ncu);
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
    }
    int bpc[plan::kNumThin];
    for (int q = 0; q < plan::kNumThin; ++q) bpc[q] = thin_occupancy(q);
    const plan::Geometry geo = plan::contract_geometry(J, m, n, pitch_s, pitch_i, r0, ncu, bpc);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    return memo.emplace(key, geo).first->second;
}

// GZ (n x k) = sig2g G S0 for the packed-genotype cohort, column c of S0 on col_limbs[c]
// int8 digits.
void mfma_geno_product(globe_ctx* c, const double* S0, int64_t ldn, int k, double* GZ, int64_t ldg,
                       const int* skip, const int8_t* col_limbs, const double* theta_cols,
                       const double* U_given, int64_t row0, const int32_t* ridx, int64_t ng) {
    hipStream_t s = c->stream;
    const int64_t n = c->n, m = c->m;
    require(k >= 1 && k <= kMaxCols, GLOBE_E_INVALID, "1 to 64 columns");
    LimbMap lm;
    std::memset(&lm, 0, sizeof(lm));
    std::memset(lm.col, 255, sizeof(lm.col));
    int J = 0;
    for (int q = 0; q < k; ++q) {
        // 0 digits: a column whose product is not needed (a frozen PCG column); its
        // result is the product of a zero vector
        require(col_limbs[q] >= 0 && col_limbs[q] <= kLimbsFP64, GLOBE_E_INVALID, "0 to 7 digits per column");
        require(J + col_limbs[q] <= 128, GLOBE_E_INVALID, "at most 128 digit columns per MFMA contraction");
        lm.L[q] = col_limbs[q];
        lm.off[q] = (uint8_t)J;
        for (int l = 0; l < col_limbs[q]; ++l) {
            lm.col[J + l] = (uint8_t)q;
            lm.dig[J + l] = (uint8_t)l;
        }
        J += col_limbs[q];
    }
    require(J >= 1, GLOBE_E_INVALID, "a contraction needs at least one digit column");
    if (c->cu_count == 0) {
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, c->device));
        c->cu_count = prop.multiProcessorCount;
    }
    const int ncu = c->cu_count;
    // pass 2 forms rows r0 .. n - 1 (r0 > 0 only with U_given), or, gathered, the ng rows
    // ridx[0 .. ng) (thin contractions only: a wide one forms every row); fewer rows can
    // take more splits, so the partial buffer is sized for the row count actually contracted
    const bool gather = ridx != nullptr && !U_given && J <= 16;
    if (gather)
        require(ng > 0 && ng % kRowPad == 0 && ng <= c->n + kRowPad, GLOBE_E_INVALID,
                "gathered rows: a positive multiple of 512, at most n + 512");
    const int64_t r0 = U_given ? row0 : 0, n2 = gather ? ng : n - r0;
    const plan::Geometry geo = product_geometry(J, m, gather ? ng : n, c->pitch_s, c->pitch_i, r0, ncu);
    const bool wide = geo.wide;
    const int JT = geo.JT, J16 = geo.J16, nfrag = geo.nfrag;
    const int64_t Kpad1 = geo.pass[0].Kpad, Kpad2 = geo.pass[1].Kpad;
    const int64_t nss1 = geo.pass[0].nss, nss2 = geo.pass[1].nss;
    // The variant of each pass: a forced candidate (validation), the tuned one, or -- in a
    // no-allocation call, which neither times candidates nor sizes the partials for all
    // of them -- the candidate with the fewest partial slots. Every candidate computes the
    // same integer sums, so the choice never changes a result.
    const bool noalloc = c->arena.base != nullptr;
    auto key_of = [&](int pass, int64_t nrows, int64_t Kpad) {
        return std::to_string(pass) + ":" + std::to_string(nrows) + ":" + std::to_string(Kpad) + ":" +
               std::to_string(J16);
    };
    auto preset = [&](int pass, int64_t nrows, int64_t Kpad, int64_t nss, Variant& v) -> bool {
        const std::vector<Variant> cand = candidates(wide, JT, nrows, nss, ncu);
        c->variant_count[pass] = (int)cand.size();
        if (c->force_variant >= 0) {
            v = cand[(size_t)c->force_variant % cand.size()];
            return true;
        }
        if (noalloc) {
            v = cand[0];
            for (const Variant& w : cand)
                if (w.S < v.S) v = w;
            return true;
        }
        auto it = c->tuned.find(key_of(pass, nrows, Kpad));
        if (it == c->tuned.end()) return false;
        v = Variant{it->second.cfg, it->second.parts, it->second.grid};
        return true;
    };
    Variant vp[2];
    bool fixed[2] = {false, false};
    if (!U_given) fixed[0] = preset(0, m, Kpad1, nss1, vp[0]);
    fixed[1] = preset(1, n2, Kpad2, nss2, vp[1]);
    if (gather && !fixed[1]) {
        // gathered rows are never timed (their count changes from call to call): the thin
        // shape tuned for all n rows, with that shape's stream-K grid nearest two rounds of
        // resident blocks
        const std::vector<Variant> cand = candidates(false, 1, n2, nss2, ncu);
        auto full = c->tuned.find(key_of(1, n, Kpad2));
        const int cfg = full != c->tuned.end() ? full->second.cfg : cand[0].cfg;
        const int64_t target = 2 * (int64_t)ncu * thin_occupancy(cfg);
        vp[1] = cand[0];
        bool have = false;
        for (const Variant& w : cand) {
            if (w.cfg != cfg) continue;
            if (!have || std::llabs(w.G - target) < std::llabs(vp[1].G - target)) vp[1] = w;
            have = true;
        }
        fixed[1] = true;
    }
    const size_t bq = (size_t)geo.bq_bytes;
    const size_t ob = noalloc ? (size_t)std::max((U_given ? 0 : vp[0].S) * m, (int64_t)vp[1].S * n2) * J16 * 4
                              : (size_t)geo.part_bytes;
    uint8_t* Bq = ws<uint8_t>(c, c->w_mix, round_up(bq, 256) + round_up(ob, 256) + 2 * kMaxCols * sizeof(int));
    int32_t* out = reinterpret_cast<int32_t*>(Bq + round_up(bq, 256));
    int* sh1 = reinterpret_cast<int*>(Bq + round_up(bq, 256) + round_up(ob, 256));
    int* sh2 = sh1 + kMaxCols;
    double* U = ws<double>(c, c->w_U, sizeof(double) * (size_t)(m > 0 ? m : 1) * k);
    double* Tc = ws<double>(c, c->w_T, sizeof(double) * kMaxCols);
    double* Cc = ws<double>(c, c->w_C, sizeof(double) * kMaxCols);
    double* part = ws<double>(c, c->w_part2, sizeof(double) * kMaxParts * kMaxCols);
    const double* th = theta_cols;
    const double flops = 2.0 * (double)n * (double)m * k;

    // run one pass with its preset variant, or tune the shape first (epi(S, skip) launches
    // the pass's epilogue over S split partials: the tuner times each candidate together
    // with it, since more splits mean more partials to combine)
    auto contract = [&](int pass, const uint8_t* rows, int64_t pitch, int64_t nrows, int64_t K,
                        int64_t Kpad, int64_t nss, const auto& epi) -> int {
        if (pass == 1 && gather) {
            // timed like a tuned launch when profiling
            globe_ctx::ProfEvent e{};
            const bool prof = c->prof_on && !noalloc;
            if (prof) {
                HIP_CHECK(hipEventCreate(&e.a));
                HIP_CHECK(hipEventCreate(&e.b));
                e.pass = pass;
                e.iter = c->prof_iter;
                e.bytes = (double)nrows * (double)((K + 3) / 4);
                e.flops = 2.0 * (double)nrows * (double)K * k;
                e.issued = 2.0 * (double)nrows * (double)Kpad * J16;
                HIP_CHECK(hipEventRecordWithFlags(e.a, s, hipEventRecordExternal));
            }
            launch_variant(s, wide, JT, vp[pass], rows, pitch, nrows, nss, Bq, out, J, skip, ridx);
            if (prof) {
                HIP_CHECK(hipEventRecordWithFlags(e.b, s, hipEventRecordExternal));
                c->prof.push_back(e);
            }
            return vp[pass].S;
        }
        if (fixed[pass] && (c->force_variant >= 0 || noalloc)) {
            launch_variant(s, wide, JT, vp[pass], rows, pitch, nrows, nss, Bq, out, J, skip);
            return vp[pass].S;
        }
        const std::string key = key_of(pass, nrows, Kpad);
        auto it = c->tuned.find(key);
        if (it == c->tuned.end() && wide && pass == 1 && r0 > 0) {
            // the GRM builder's rows r0 .. n - 1 (r0 a multiple of 512): the block shape
            // tuned for all n rows, with that shape's candidate split for the shorter
            // range nearest the tuned one (fewer is taken on a tie: fewer partials to
            // combine), so each row range is not tuned anew
            auto full = c->tuned.find(key_of(pass, n, Kpad));
            if (full != c->tuned.end()) {
                const int cfg = full->second.cfg, S_full = full->second.parts;
                int S = 0;
                for (const Variant& w : candidates(wide, JT, nrows, nss, ncu))
                    if (w.cfg == cfg && (S == 0 || std::abs(std::log2((double)w.S / S_full)) <
                                                       std::abs(std::log2((double)S / S_full))))
                        S = w.S;
                if (S > 0) it = c->tuned.emplace(key, globe_ctx::Tuned{cfg, S, 0}).first;
            }
        }
        if (it == c->tuned.end()) {
            const std::vector<Variant> cand = candidates(wide, JT, nrows, nss, ncu);
            hipEvent_t e0, e1;
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            float best = 1e30f;
            int bi = 0;
            for (size_t q = 0; q < cand.size(); ++q) {
                HIP_CHECK(hipEventRecord(e0, s));
                launch_variant(s, wide, JT, cand[q], rows, pitch, nrows, nss, Bq, out, J, nullptr);
                epi(cand[q].S, nullptr);
                HIP_CHECK(hipEventRecord(e1, s));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (std::getenv("GLOBE_TUNE_LOG"))
                    std::fprintf(stderr, "[globe tune] %s shape %d parts %d grid %lld: %.3f ms\n",
                                 key.c_str(), cand[q].cfg, cand[q].S, (long long)cand[q].G, ms);
                if (ms < best) {
                    best = ms;
                    bi = (int)q;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            it = c->tuned.emplace(key, globe_ctx::Tuned{cand[bi].cfg, cand[bi].S, cand[bi].G}).first;
        }
        const Variant v{it->second.cfg, it->second.parts, it->second.grid};
        if (c->prof_on && !noalloc) {
            globe_ctx::ProfEvent e{};
            HIP_CHECK(hipEventCreate(&e.a));
            HIP_CHECK(hipEventCreate(&e.b));
            e.pass = pass;
            e.iter = c->prof_iter;
            e.bytes = (double)nrows * (double)((K + 3) / 4);
            e.flops = flops;
            e.issued = 2.0 * (double)nrows * (double)Kpad * J16;
            // external: inside a captured PCG iteration the record is a graph node that
            // stamps the event on every replay
            HIP_CHECK(hipEventRecordWithFlags(e.a, s, hipEventRecordExternal));
            launch_variant(s, wide, JT, v, rows, pitch, nrows, nss, Bq, out, J, skip);
            HIP_CHECK(hipEventRecordWithFlags(e.b, s, hipEventRecordExternal));
            c->prof.push_back(e);
        } else {
            launch_variant(s, wide, JT, v, rows, pitch, nrows, nss, Bq, out, J, skip);
        }
        return v.S;
    };

    int64_t eb1 = 0;
    if (U_given) {
        // pass 1 done by the caller: u = U_given (m x k); its column maxima set the scales
        U = const_cast<double*>(U_given);
        eb1 = parts_for(m);
        k_colmax_parts<<<dim3((unsigned)eb1, k), 256, 0, s>>>(m, U, m, part, skip);
        HIP_CHECK(hipGetLastError());
    } else {
    // pass 1: P1[s][c] = sum_i g_is S0[c][i] on the SNP-major rows
    const int npn = parts_for(n);
    k_colmax_parts<<<dim3(npn, k), 256, 0, s>>>(n, S0, ldn, part, skip);
    k_finish_cols<<<1, kFinishThreads, 0, s>>>(part, npn, k, 1, lm, sh1, nullptr, skip);
    const int qg1 = quantize_parts(Kpad1);
    k_quantize<<<dim3(qg1, k + 1), 256, 0, s>>>(S0, ldn, n, Kpad1, k, lm, nfrag, wide, J, J16, sh1, nullptr,
                                               Bq, part, skip);
    k_finish_cols<<<1, kFinishThreads, 0, s>>>(part, qg1, k, 0, lm, nullptr, Tc, skip);
    HIP_CHECK(hipGetLastError());
    eb1 = std::min<int64_t>((m + kEpiRows - 1) / kEpiRows, kMaxParts);
    auto epi1 = [&](int S, const int* sk) {
        k_epilogue<<<(unsigned)eb1, 256, 0, s>>>(out, S, m, J16, lm, k, sh1, 0, Tc,
                                                 c->d_twof.as<double>(), c->d_inv.as<double>(), th,
                                                 1.0 / (double)m, U, m, part, sk, nullptr);
        HIP_CHECK(hipGetLastError());
    };
    if (gather) {
        // S0 is zero outside the listed individuals: contract over their rows of the
        // individual-major layout (k_geno_tg) instead of every individual's column of the
        // SNP-major one. The digits are k_quantize's for those individuals (k_quantize
        // above still forms T, the column sums); Bq's pass-1 digits are not read, so the
        // listed individuals' digits take its place. Same integer sums, same epilogue.
        require(J16 == 16 && (size_t)16 * (size_t)ng <= bq, GLOBE_E_INVALID, "gathered pass 1 shape");
        int8_t* Dt = reinterpret_cast<int8_t*>(Bq);
        k_quantize_gather<<<dim3((unsigned)std::min<int64_t>((ng + 255) / 256, 1024), 16), 256, 0, s>>>(
            S0, ldn, ridx, ng, lm, J, sh1, Dt, ng, skip);
        HIP_CHECK(hipGetLastError());
        constexpr int kRB = 256;  // row bytes per workgroup
        const plan::TgLaunch tl = plan::tg_launch<kRB>(c->pitch_i, m, ng, ncu, (int64_t)ob);
        const int64_t R = tl.R;
        const int P = tl.P;
        globe_ctx::ProfEvent e{};
        const bool prof = c->prof_on && !noalloc;
        if (prof) {
            HIP_CHECK(hipEventCreate(&e.a));
            HIP_CHECK(hipEventCreate(&e.b));
            e.pass = 0;
            e.iter = c->prof_iter;
            e.bytes = (double)ng * (double)((m + 3) / 4);
            e.flops = 2.0 * (double)ng * (double)m * k;
            e.issued = 2.0 * (double)ng * (double)(4 * c->pitch_i) * J16;
            HIP_CHECK(hipEventRecordWithFlags(e.a, s, hipEventRecordExternal));
        }
        k_geno_tg<kRB><<<dim3((unsigned)R, (unsigned)P), 512, 0, s>>>(c->d_ind.as<uint8_t>(), c->pitch_i, m, ridx,
                                                                     ng, Dt, ng, P, out, skip);
        HIP_CHECK(hipGetLastError());
        if (prof) {
            HIP_CHECK(hipEventRecordWithFlags(e.b, s, hipEventRecordExternal));
            c->prof.push_back(e);
        }
        epi1(P, skip);
    } else {
        epi1(contract(0, c->d_snp.as<uint8_t>(), c->pitch_s, m, n, Kpad1, nss1, epi1), skip);
    }
    }
    // pass 2: GZ[c][i] = sum_s g_is u[c][s] - C[c] on the individual-major rows
    k_finish_cols<<<1, kFinishThreads, 0, s>>>(part, (int)eb1, k, 1, lm, sh2, nullptr, skip);
    const int qg2 = quantize_parts(Kpad2);
    k_quantize<<<dim3(qg2, k + 1), 256, 0, s>>>(U, m, m, Kpad2, k, lm, nfrag, wide, J, J16, sh2,
                                               c->d_twof.as<double>(), Bq, part, skip);
    k_finish_cols<<<1, kFinishThreads, 0, s>>>(part, qg2, k, 0, lm, nullptr, Cc, skip);
    HIP_CHECK(hipGetLastError());
    const int64_t eb2 = std::min<int64_t>((n2 + kEpiRows - 1) / kEpiRows, kMaxParts);
    auto epi2 = [&](int S, const int* sk) {
        k_epilogue<<<(unsigned)eb2, 256, 0, s>>>(out, S, n2, J16, lm, k, sh2, 1, Cc, nullptr,
                                                 nullptr, nullptr, 0.0, GZ + r0, ldg, nullptr, sk,
                                                 gather ? ridx : nullptr);
        HIP_CHECK(hipGetLastError());
    };
    epi2(contract(1, c->d_ind.as<uint8_t>() + r0 * c->pitch_i, c->pitch_i, n2, m, Kpad2, nss2, epi2), skip);
}

}  // namespace globe

namespace globe {

// Tune the contraction for every digit-column width class (thin, 1 to 4 wide fragments of
// 32) that a block of up to max_cols columns can reach at `precision`, so that no shape is
// first tuned inside a running solve.
void pretune_geno_product(globe_ctx* c, int max_cols, int precision, const double* theta_cols) {
    if (c->gen != 1 || c->m == 0 || precision == GLOBE_PREC_FP64_VALU) return;
    const int L = precision == GLOBE_PREC_FP64 ? kLimbsFP64 : kLimbsMixed;
    const int maxJ = std::min(128, L * std::max(1, max_cols));
    const int64_t n = c->n;
    const int kmax = (128 + L - 1) / L;
    c->w_P2.ensure(sizeof(double) * (size_t)n * kmax);
    c->w_P1.ensure(sizeof(double) * (size_t)n * kmax);
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n * kmax + 255) / 256, 2048));
    k_tune_fill<<<gx, 256, 0, c->stream>>>(c->w_P2.as<double>(), n * kmax, 0x5EED5EEDull);
    HIP_CHECK(hipGetLastError());
    // classes of J: [1, 16] thin, then 1, 2, 3, 4 wide fragments
    const int kClass[5][2] = {{1, 16}, {17, 32}, {33, 64}, {65, 96}, {97, 128}};
    for (const auto& cl : kClass) {
        const int lo = cl[0], hi = cl[1];
        if (lo > maxJ) break;
        const int J = std::min(hi, maxJ);
        const int kk = (J + L - 1) / L;
        std::vector<int8_t> limbs((size_t)kk, (int8_t)L);
        limbs[(size_t)kk - 1] = (int8_t)(J - L * (kk - 1));
        mfma_geno_product(c, c->w_P2.as<double>(), n, kk, c->w_P1.as<double>(), n, nullptr, limbs.data(),
                          theta_cols);
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

// Rows row0 .. n - 1 of GZ (n x k) = G [e_j0 .. e_j0+k-1] for the GRM builder: pass 1 is
// the direct row extraction k_unit_u, pass 2 the contraction on kBuilderLimbs digits.
// u = W^T e_j / M has no large common offset (its entries inv_s^2 (g_sj - 2 f_s) / M
// change sign with g_sj), so G_ij = sum_s g_is u_s - C is not a difference of large
// terms. 4 digits hold u to 2^-30 of its column maximum (|u| <= 2 inv_s^2 / M, at most
// 101 / M at MAF 0.01): the M rounding errors of an entry add as a random walk, about
// 1e-10, and even all aligned (2 M 2^-31 max|u| <= 1e-7) stay at the entry's FP32
// rounding. 32 columns of 4 digits fill one 128-column contraction, and 4-digit columns
// take the epilogue's vector path.
void geno_unit_columns(globe_ctx* c, int64_t j0, int k, double* GZ, int64_t row0) {
    constexpr int kBuilderLimbs = 4;
    hipStream_t s = c->stream;
    const int64_t m = c->m;
    c->w_P1.ensure(sizeof(double) * (size_t)(m > 0 ? m : 1) * k);
    double* U = c->w_P1.as<double>();
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, 1024));
    k_unit_u<<<dim3(gx, k), 256, 0, s>>>(c->d_ind.as<uint8_t>(), c->pitch_i, j0, m, c->d_twof.as<double>(),
                                         c->d_inv.as<double>(), 1.0 / (double)m, U);
    HIP_CHECK(hipGetLastError());
    const int kg = 128 / kBuilderLimbs;
    for (int c0 = 0; c0 < k; c0 += kg) {
        const int kk = std::min(kg, k - c0);
        const std::vector<int8_t> L((size_t)kk, (int8_t)kBuilderLimbs);
        mfma_geno_product(c, nullptr, c->n, kk, GZ + (size_t)c0 * c->n, c->n, nullptr, L.data(), nullptr,
                          U + (size_t)c0 * m, row0);
    }
}

}  // namespace globe
