// Launch plans of the packed-genotype contraction (k_mfma.hip): block shapes, split-K and
// stream-K decompositions, the tuner's candidate list, and the index arithmetic every
// kernel uses for its global reads and partial-sum stores.
//
// The kernels and the host replay (plan_replay.cpp, which walks every candidate's blocks
// on the CPU and checks each read and store against its buffer) both call these
// functions, so the replay checks the arithmetic the GPU executes rather than a copy of it.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

namespace globe {
namespace plan {

// ---------------------------------------------------------------- block shapes
// wide kernel (k_geno_wide): waves, ring stages, column groups, row fragments per wave,
// PF: 1 reads the digits from LDS two K-steps ahead of their MFMAs, pinned in that
// order (0: one K-step ahead, as the compiler schedules; PF changes only the order of
// the LDS reads), and KS: K-blocks of 128 genotypes per stage. KS = 2 stages a whole
// super-step (64 packed bytes of a row per copy instruction instead of 32: the row
// stream reads each 128-byte line in two requests rather than four) and halves the
// barriers per MFMA.
struct WideShape {
    int W, NS, CG, RA, PF, KS;
};
constexpr int kNumWide = 14;
constexpr WideShape kWide[kNumWide] = {
    {8, 4, 1, 2, 0, 1}, {4, 3, 1, 2, 0, 1}, {8, 4, 2, 4, 0, 1}, {4, 3, 2, 4, 0, 1}, {8, 4, 1, 2, 1, 1},
    {4, 3, 1, 2, 1, 1}, {8, 4, 2, 4, 1, 1}, {4, 3, 2, 4, 1, 1}, {4, 5, 1, 2, 1, 1}, {8, 6, 1, 2, 1, 1},
    {8, 2, 1, 2, 1, 2}, {8, 2, 2, 4, 1, 2}, {4, 2, 1, 2, 1, 2}, {8, 3, 1, 2, 1, 2}};
// thin kernel (k_geno_thin): waves, row fragments of 16 rows per wave, ring stages,
// super-steps (256 genotypes = 64 packed bytes of a row) per stage
struct ThinShape {
    int W, RT, NS, SS;
};
constexpr int kNumThin = 8;
constexpr ThinShape kThin[kNumThin] = {{4, 4, 3, 1},  {8, 2, 3, 2},  {8, 2, 4, 2}, {16, 1, 3, 2},
                                       {16, 1, 4, 2}, {8, 1, 5, 2}, {4, 4, 4, 2}, {8, 1, 6, 2}};

// Per-stage layout of the wide kernel <JT column fragments of 32 digit columns, W waves,
// CG column groups, RA row fragments of 32 rows per wave, KS K-blocks of 128 genotypes
// per stage>: the waves form W / CG row groups x CG column groups; a wave owns RA row
// fragments and CB = JT / CG column fragments, so a block covers (W / CG) 32 RA rows x all
// 32 JT digit columns. A row fragment takes KS KB of a stage (32 rows x 32 KS packed
// bytes), the digits 4 JT KS chunks of 1 KB (K-step major, then column fragment).
template <int JT, int W, int CG, int RA, int KS = 1>
struct WideCfg {
    static constexpr int kW = W, kJT = JT, kRA = RA, kKS = KS;
    static constexpr int kCB = JT / CG;                 // column fragments per wave
    static constexpr int kRowGroups = W / CG;
    static constexpr int kRows = kRowGroups * 32 * RA;  // rows per block
    static constexpr int kAFrags = kRowGroups * RA;     // row fragments per stage
    static constexpr int kAPerWave = kAFrags / W;       // row fragments copied per wave
    static constexpr int kA = kAFrags * 1024 * KS;
    static constexpr int kB = 4 * JT * 1024 * KS;       // digits of a stage
    static constexpr int kSlot = kA + kB;
    static constexpr int kBChunks = 4 * JT * KS;        // 1 KB digit chunks
    static constexpr int kBPerWave = kBChunks >= W ? kBChunks / W : 1;
    static constexpr int kSink = kBChunks < W ? (W - kBChunks) * 1024 : 0;
    static constexpr int kOps = KS * kAPerWave + kBPerWave;  // copies per wave per stage
    static_assert(JT % CG == 0 && W % CG == 0, "column groups");
    static_assert(kAFrags % W == 0, "row fragments split evenly over waves");
    static_assert(kBChunks < W || kBChunks % W == 0, "digit chunks split evenly over waves");
    static_assert(KS == 1 || KS == 2, "one or two K-blocks per stage");
};

// Per-stage layout of the thin kernel <W waves, RT row fragments of 16 rows per wave, SS
// super-steps per stage> (one column fragment of 16 digit columns).
template <int W, int RT, int SS>
struct ThinCfg {
    static constexpr int kRB = 16 * W * RT;              // rows per row block
    static constexpr int kRowKB = W * RT * SS;           // packed rows of a stage
    static constexpr int kDigKB = 4 * SS;                // digits of a stage (4 K-steps per super-step)
    static constexpr int kA = kRowKB * 1024;
    static constexpr int kB = kDigKB * 1024;
    static constexpr int kSlot = kA + kB;
    // digit chunks per wave: spread evenly, or one each for waves < kDigKB when there
    // are more waves than chunks (the other waves copy only rows)
    static constexpr int kDigPerWave = kDigKB >= W ? kDigKB / W : 1;
    static constexpr int kOps = RT * SS + kDigPerWave;  // copies per stage, digit waves
    static constexpr int kOpsRowOnly = RT * SS;         // copies per stage, other waves
    static_assert(W == 4 || W == 8 || W == 16, "4, 8 or 16 waves");
    static_assert(SS == 1 || SS == 2, "1 or 2 super-steps per stage");
    static_assert(kDigKB < W || kDigKB % W == 0, "digit chunks must spread evenly");
};

// A wide shape fits JT column fragments when the column groups divide them, the stage's
// 4 JT KS digit chunks spread evenly over the waves (or number fewer than them) and its
// ring of NS stages fits the 160 KB of LDS.
constexpr bool wide_fits(int JT, int q) {
    return JT % kWide[q].CG == 0 &&
           (4 * JT * kWide[q].KS < kWide[q].W || (4 * JT * kWide[q].KS) % kWide[q].W == 0) &&
           kWide[q].NS * ((kWide[q].W / kWide[q].CG) * kWide[q].RA + 4 * JT) * kWide[q].KS * 1024 +
                   (4 * JT * kWide[q].KS < kWide[q].W ? (kWide[q].W - 4 * JT * kWide[q].KS) * 1024 : 0) <=
               160 * 1024;
}

inline int64_t rows_per_block(bool wide, int cfg) {
    if (!wide) return (int64_t)16 * kThin[cfg].W * kThin[cfg].RT;
    const WideShape& w = kWide[cfg];
    return (int64_t)(w.W / w.CG) * 32 * w.RA;
}

// ---------------------------------------------------------------- wide: split-K
// Split-K is a power of two S; block b works on K-split b % S of row block b / S for
// S <= 8, and for S > 8 every XCD (b % 8) walks its splits one after another
// (split = b % 8 + 8 ((b / 8) / R)), so the blocks resident on one XCD share one K-chunk
// of the digit matrix in that XCD's L2.
__host__ __device__ inline void block_coords(int64_t b, int64_t R, int S, int64_t& rb, int& sp) {
    if (S <= 8) {
        sp = (int)(b % S);
        rb = b / S;
    } else {
        const int64_t t = b / 8;
        sp = (int)(b % 8) + 8 * (int)(t / R);
        rb = t % R;
    }
}

// Block b of a wide launch: its row block and split, first row, first stage H0 (128
// genotypes each) and stage count nt of its K range (super-steps [ss0, ss1) of 256).
struct WideBlock {
    int64_t rb, blk_row, H0, nt;
    int sp;
};
__host__ __device__ inline WideBlock wide_block(int64_t b, int64_t R, int S, int64_t rows_per_blk,
                                                int64_t n_ss, int64_t ss_per_split) {
    WideBlock w;
    block_coords(b, R, S, w.rb, w.sp);
    w.blk_row = w.rb * rows_per_blk;
    const int64_t ss0 = (int64_t)w.sp * ss_per_split;
    const int64_t ss1 = ss0 + ss_per_split < n_ss ? ss0 + ss_per_split : n_ss;
    w.H0 = 2 * ss0;
    w.nt = ss1 > ss0 ? 2 * (ss1 - ss0) : 0;
    return w;
}
// Packed-row reads of a wide block: lane (rr = lane % 32, g = lane / 32) copies the 16
// bytes at wide_row_base + wide_stage_step(t) + f wide_frag_stride for row fragment f
// (rows blk_row + 32 f ..) at stage H0 + t (32 packed bytes per row and stage).
__host__ __device__ inline int64_t wide_row_base(int64_t blk_row, int rr, int g, int64_t H0, int64_t pitch) {
    return (blk_row + rr) * pitch + 16 * g + 32 * H0;
}
__host__ __device__ inline int64_t wide_stage_step(int64_t t) { return 32 * t; }
__host__ __device__ inline int64_t wide_frag_stride(int64_t pitch) { return 32 * pitch; }
// Packed-row reads with 256-genotype stages (KS = 2): row fragment f of a stage holds 32
// rows x 64 packed bytes in 2 KB of LDS, row r at 64 r with its 16-byte chunk c at
// position c ^ wide_swz(r), so that the MFMA reads of one chunk by rows 0..31 spread over
// the LDS banks. Copy instruction i (0, 1) of a fragment: lane -> row 16 i + lane / 4, LDS
// position lane % 4, i.e. every row's 4 lanes read its 64 contiguous bytes and the 64
// lanes land as one contiguous KB. Offsets are from the block's first row at its first
// stage, wide_row_base(blk_row, 0, 0, H0, pitch); t counts 256-genotype stages.
__host__ __device__ constexpr int wide_swz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
__host__ __device__ inline int wide2_copy_row(int i, int lane) { return 16 * i + (lane >> 2); }
__host__ __device__ inline int wide2_copy_chunk(int i, int lane) {
    return (lane & 3) ^ wide_swz(wide2_copy_row(i, lane));
}
__host__ __device__ inline int64_t wide2_lane_off(int f, int i, int lane, int64_t pitch) {
    return (int64_t)(32 * f + wide2_copy_row(i, lane)) * pitch + 16 * wide2_copy_chunk(i, lane);
}
__host__ __device__ inline int64_t wide2_stage_step(int64_t t) { return 64 * t; }
// LDS byte of row rr's chunk c in a 2 KB fragment
__host__ __device__ inline int wide2_lds_chunk(int rr, int c) { return rr * 64 + ((c ^ wide_swz(rr)) << 4); }
// Digit reads: lane copies the 16 bytes at dig_stage(H, kB) + dig_lane(chunk, lane) from
// 1 KB chunk `chunk` of stage H (kB bytes of digits per stage; both kernels).
__host__ __device__ inline int64_t dig_stage(int64_t H, int kB) { return H * kB; }
__host__ __device__ inline int dig_lane(int chunk, int lane) { return chunk * 1024 + lane * 16; }
// Row of accumulator element v (0..15) of row fragment a, for lane half g, of the wave
// whose row group starts at row_base (32x32 MFMA output layout).
__host__ __device__ inline int64_t wide_out_row(int64_t row_base, int a, int v, int g) {
    return row_base + a * 32 + 8 * (v >> 2) + 4 * g + (v & 3);
}
// int32 partial-sum slot of (split sp, row r, digit column col): out[sp][row][J16]
__host__ __device__ inline int64_t part_idx(int64_t sp, int64_t nrows, int64_t r, int J16, int col) {
    return (sp * nrows + r) * J16 + col;
}

// ---------------------------------------------------------------- thin: stream-K
// Units u = r n_st + t (row block r, stage t of its n_st stages), u < U = R n_st: block b
// takes units [b U / G, (b + 1) U / G) as one continuous ring of stages. A row block is
// covered by consecutive blocks: the one holding its stage 0 writes partial slot 0, the
// next slot 1, ... (slot = b - first block of the row block), and the block that ends a
// row block zeroes the slots above its own up to P - 1.
__host__ __device__ inline int64_t sk_block_of(int64_t u, int64_t U, int64_t G) {
    return ((u + 1) * G - 1) / U;  // the b with b U / G <= u < (b + 1) U / G
}
__host__ __device__ inline int64_t sk_lo(int64_t b, int64_t U, int64_t G) { return b * U / G; }
__host__ __device__ inline int64_t sk_slot(int64_t b, int64_t cr, int64_t n_st, int64_t U, int64_t G) {
    return b - sk_block_of(cr * n_st, U, G);
}
// Packed-row reads of a thin block: lane (rr = lane % 16, g = lane / 16) of wave `wave`
// copies the 16 bytes at thin_unit_off(ir, it) + thin_lane_off + thin_frag_off(q, h) for
// row fragment q, super-step h of stage it of row block ir (64 SS packed bytes per row and
// stage).
__host__ __device__ inline int64_t thin_unit_off(int64_t ir, int64_t kRB, int64_t it, int SS, int64_t pitch) {
    return ir * kRB * pitch + 64 * SS * it;
}
__host__ __device__ inline int64_t thin_lane_off(int wave, int RT, int rr, int g, int64_t pitch) {
    return (int64_t)(wave * 16 * RT + rr) * pitch + 16 * g;
}
__host__ __device__ inline int64_t thin_frag_off(int q, int h, int64_t pitch) { return q * (16 * pitch) + 64 * h; }
// Row of accumulator element v (0..3) of row fragment a for lane group g (16x16 MFMA
// output layout) in row block cr.
__host__ __device__ inline int64_t thin_out_row(int64_t cr, int64_t kRB, int wave, int RT, int a, int v, int g) {
    return cr * kRB + wave * 16 * RT + 4 * g + a * 16 + v;
}

// Copy cursor of a stream-K block over its units [lo, lo + nt): (row block r, stage t) of
// the next copy. It stops at the block's last unit, so the ring's last NS - 1 copies
// repeat that unit into slots that are never multiplied.
struct SkCursor {
    int64_t r, t, issued, nt, n_st;
    __host__ __device__ SkCursor(int64_t lo, int64_t nt_, int64_t n_st_)
        : r(lo / n_st_), t(lo % n_st_), issued(0), nt(nt_), n_st(n_st_) {}
    __host__ __device__ void advance() {
        if (issued + 1 < nt) {
            ++issued;
            if (++t == n_st) {
                t = 0;
                ++r;
            }
        }
    }
};

// Pieces per row block of the stream-K split: the most blocks any row block spans.
inline int thin_pieces(int64_t R, int64_t nst, int64_t G) {
    const int64_t U = R * nst;
    int64_t P = 1;
    for (int64_t r = 0; r < R; ++r)
        P = std::max(P, sk_block_of(r * nst + nst - 1, U, G) - sk_block_of(r * nst, U, G) + 1);
    return (int)P;
}

// ---------------------------------------------------------------- variants
// Split-K: enough blocks to fill the chip (>= 4 per CU), each split >= 8 super-steps;
// wide contractions split at least 8 ways so each XCD streams one K-chunk of digits.
inline int choose_split(int64_t R, int64_t n_ss, bool wide) {
    int S = 1;
    while (R * S < 1024 && S < 64) S <<= 1;
    if (wide && S < 8) S = 8;
    while (S > 1 && n_ss / S < 8) S >>= 1;
    return S;
}

struct Variant {
    int cfg;     // index into kThin (thin) or kWide (wide)
    int S;       // partial-sum slots: split-K (wide), pieces per row block (thin)
    int64_t G;   // thin: stream-K grid (blocks); wide: 0
};

// The tuner's candidates: the thin block shapes of kThin with stream-K grids that fill
// the chip in 1, 2 or 4 rounds of resident blocks or take 1-2 blocks per row block; the
// wide block shapes of kWide that fit JT, each with split-K S/2, S, 2 S around
// choose_split(). thin_bpc[cfg]: resident blocks per CU of thin shape cfg.
inline std::vector<Variant> candidates(bool wide, int JT, int64_t nrows, int64_t nss, int ncu,
                                       const int* thin_bpc) {
    std::vector<Variant> v;
    if (wide) {
        for (int cfg = 0; cfg < kNumWide; ++cfg) {
            if (!wide_fits(JT, cfg)) continue;
            const int64_t R = (nrows + rows_per_block(true, cfg) - 1) / rows_per_block(true, cfg);
            const int S0 = choose_split(R, nss, true);
            for (int S : {S0 / 2, S0, S0 * 2}) {
                if (S < 1 || S > 64 || (S > 1 && nss / S < 4)) continue;
                v.push_back({cfg, S, 0});
            }
        }
        return v;
    }
    for (int cfg = 0; cfg < kNumThin; ++cfg) {
        if (nss % kThin[cfg].SS != 0) continue;
        const int64_t R = (nrows + rows_per_block(false, cfg) - 1) / rows_per_block(false, cfg);
        const int64_t U = R * (nss / kThin[cfg].SS);
        const int64_t resident = (int64_t)ncu * std::max(1, thin_bpc[cfg]);
        for (int64_t G : {resident, 2 * resident, 4 * resident, R, 2 * R}) {
            if (G < 1 || G > U) continue;
            const int P = thin_pieces(R, nss / kThin[cfg].SS, G);
            if (P > 64) continue;
            bool dup = false;
            for (const Variant& w : v) dup |= w.cfg == cfg && w.G == G;
            if (!dup) v.push_back({cfg, P, G});
        }
    }
    return v;
}

// Grid of a variant: R row blocks x S splits (wide) or the stream-K grid G (thin).
inline int64_t variant_rowblocks(bool wide, const Variant& v, int64_t nrows) {
    return (nrows + rows_per_block(wide, v.cfg) - 1) / rows_per_block(wide, v.cfg);
}
inline int64_t variant_grid(bool wide, const Variant& v, int64_t nrows) {
    return wide ? variant_rowblocks(true, v, nrows) * v.S : v.G;
}

// Digit matrix layout: digit column j (= fragment jf of fw = 16 (thin) or 32 (wide)
// columns, column jj in it) of genotype kap sits at byte bq_index of Bq, so that lane l
// of a K-step's 1 KB chunk holds 16 consecutive digits of one column in the decode order
// {0,4,8,12, 1,5,9,13, 2,6,10,14, 3,7,11,15} of the packed dword.
__host__ __device__ inline int64_t bq_index(int64_t kap, int jj, int jf, int nfrag, bool wide) {
    const int64_t ss = kap >> 8;
    const int rem = (int)(kap & 255);
    const int w = rem & 15;
    const int qslot = 4 * (w & 3) + (w >> 2);
    int64_t ks;
    int ln;
    if (wide) {
        // stage 2 ss + rem / 128; within it lane half (rem / 64) % 2, K-step (rem / 16) % 4
        ks = 8 * ss + 4 * (rem >> 7) + ((rem >> 4) & 3);
        ln = 32 * ((rem >> 6) & 1) + jj;
    } else {
        ks = 4 * ss + ((rem >> 4) & 3);
        ln = 16 * (rem >> 6) + jj;
    }
    return ((ks * nfrag + jf) * 64 + ln) * 16 + qslot;
}

// ---------------------------------------------------------------- gathered pass 1
// k_geno_tg: pass 1 over the rows of ng listed individuals of the individual-major layout.
// Workgroup (x, sp) owns bytes RB x .. RB x + RB - 1 of every row (4 RB SNPs) and split sp
// of the ng / 64 stages of 64 listed individuals; its 8 waves copy each stage's 64 rows x
// RB bytes into LDS with 1 KB copy instructions (RB / 128 per wave), 16-byte pieces
// swizzled by row.
template <int RB> __host__ __device__ constexpr int tg_pieces() { return RB / 16; }
template <int RB> __host__ __device__ constexpr int tg_copies() { return RB / 128; }
// LDS slot of piece p in row r is p ^ tg_swz(r): 128-byte rows put rows r and r + 2 in the
// same banks, 256-byte rows every row
template <int RB> __host__ __device__ inline int tg_swz(int r) { return RB == 128 ? ((r >> 1) & 7) : (r & 7); }
// stage row (0..63) filled by copy instruction i of wave w, lane `lane`
template <int RB> __host__ __device__ inline int tg_copy_row(int i, int w, int lane) {
    return (8 * i + w) * (1024 / RB) + lane / tg_pieces<RB>();
}
// byte offset inside its row of the global piece that lane `lane` copies for stage row r
// (the lane's LDS slot is lane % pieces, so the piece is that slot's swizzle partner),
// clamped inside the pitch for the last workgroup of a pitch that RB does not divide
template <int RB> __host__ __device__ inline int64_t tg_piece_off(int64_t b0, int r, int lane, int64_t pitch) {
    const int64_t off = b0 + 16 * ((lane % tg_pieces<RB>()) ^ tg_swz<RB>(r));
    return off + 16 <= pitch ? off : pitch - 16;
}
// LDS byte of logical piece p of stage row r
template <int RB> __host__ __device__ inline int tg_row_slot(int r, int p) {
    return r * RB + ((p ^ tg_swz<RB>(r)) << 4);
}
// digits: lanes 0..7 of wave w copy digit column n = 2 w + lane / 4, 16-byte chunk
// g = (lane & 3) ^ ((n >> 2) & 3) of a stage's 64 individuals (column n at n ldd)
__host__ __device__ inline int tg_dig_col(int w, int lane) { return 2 * w + (lane >> 2); }
__host__ __device__ inline int64_t tg_dig_off(int w, int lane, int64_t ldd) {
    const int n = tg_dig_col(w, lane);
    return (int64_t)n * ldd + 16 * ((lane & 3) ^ ((n >> 2) & 3));
}
// LDS byte of digit chunk g of column n
__host__ __device__ inline int tg_dig_slot(int n, int g) { return n * 64 + ((g ^ ((n >> 2) & 3)) << 4); }
// first stage of split sp of nst stages in nsplit splits
__host__ __device__ inline int64_t tg_split_lo(int64_t sp, int64_t nst, int64_t nsplit) { return sp * nst / nsplit; }
// SNP of accumulator element v of MFMA group g, copy column i of wave w, field j
template <int RB> __host__ __device__ inline int64_t tg_out_snp(int64_t b0, int w, int i, int g, int v, int j) {
    return 4 * (b0 + 16 * (tg_copies<RB>() * w + i) + 4 * g + v) + j;
}
// Launch: R workgroups across the pitch and P splits of the listed individuals -- four
// workgroups per CU, each split at least 16 stages long, at most 8 splits and at most
// the slots of m x 16 int32 partials that the partial buffer holds (part_bytes)
struct TgLaunch {
    int64_t R;
    int P;
};
template <int RB> inline TgLaunch tg_launch(int64_t pitch, int64_t m, int64_t ng, int ncu, int64_t part_bytes) {
    const int64_t R = (pitch + RB - 1) / RB;
    const int64_t pmax = std::max<int64_t>(1, part_bytes / (m * 16 * (int64_t)sizeof(int32_t)));
    const int64_t P = std::max<int64_t>(1, std::min({(4 * (int64_t)ncu + R - 1) / R, ng / 64 / 16, pmax, (int64_t)8}));
    return {R, (int)P};
}

// ---------------------------------------------------------------- geometry of a product
// One contraction pass: `nrows` rows of `pitch` packed bytes (the launch's row pointer
// may start inside the layout), Kpad genotypes per padded row, nss super-steps of 256.
struct PassGeom {
    int64_t nrows, pitch, Kpad, nss;
};
// Both passes of G S0 on J digit columns (W^T z over the m SNP-major rows, then W u over
// the individual-major rows r0 .. n - 1), and the scratch they share: the digit matrix
// (the larger pass's) and the int32 split partials (the most slots any candidate of
// either pass writes, for its row count).
struct Geometry {
    bool wide;
    int JT, J16, nfrag;
    PassGeom pass[2];
    int64_t bq_bytes, part_bytes;
};
inline Geometry contract_geometry(int J, int64_t m, int64_t n, int64_t pitch_s, int64_t pitch_i, int64_t r0,
                                  int ncu, const int* thin_bpc) {
    Geometry g;
    g.wide = J > 16;
    g.JT = g.wide ? (J + 31) / 32 : 1;  // column fragments of 32 digit columns
    g.J16 = g.wide ? 32 * g.JT : 16;
    g.nfrag = g.wide ? g.JT : 1;
    g.pass[0] = {m, pitch_s, pitch_s * 4, pitch_s * 4 / 256};
    g.pass[1] = {n - r0, pitch_i, pitch_i * 4, pitch_i * 4 / 256};
    g.bq_bytes = std::max(g.pass[0].Kpad, g.pass[1].Kpad) * g.J16;
    int64_t slots_rows = 0;
    for (int q = 0; q < 2; ++q) {
        int smax = 1;
        for (const Variant& v : candidates(g.wide, g.JT, g.pass[q].nrows, g.pass[q].nss, ncu, thin_bpc))
            smax = std::max(smax, v.S);
        slots_rows = std::max(slots_rows, (int64_t)smax * g.pass[q].nrows);
    }
    g.part_bytes = slots_rows * g.J16 * 4;
    return g;
}

}  // namespace plan
}  // namespace globe
