// Host replay of the packed-genotype contraction's launch plans (libglobe_plan_replay.so,
// a checker: no device code, no GPU needed).
//
// For one product shape -- J digit columns, m SNP-major rows, n individual-major rows, the
// builder's first row r0 -- it takes every candidate variant of both passes exactly as
// the tuner lists them (plan::candidates) and walks every block of each launch through
// the index functions the kernels call (contract_plan.h): every global read of packed
// rows must fall inside the padded layout, every digit read and quantize write inside
// the digit matrix, every partial-sum store and epilogue read inside the split-partial
// buffer that mfma_geno_product allocates (plan::contract_geometry), and every partial
// slot the epilogue adds must be written exactly once by the launch (split-K: the (row
// block, split) pairs and K ranges tile exactly; stream-K: each row block's pieces fill
// slots 0 .. P - 1). Row offsets are monotone in every lane, wave and fragment index, so
// the extreme lanes of each copy bound all of its lanes.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "contract_plan.h"
#include "globe_internal.h"

namespace {

using namespace globe;
using namespace globe::plan;

struct Fail {
    std::string msg;
};

void check(bool ok, const std::string& what) {
    if (!ok) throw Fail{what};
}

std::string vname(int pass, bool wide, const Variant& v) {
    char b[160];
    std::snprintf(b, sizeof b, "pass %d %s shape %d slots %d grid %lld", pass + 1, wide ? "wide" : "thin", v.cfg,
                  v.S, (long long)v.G);
    return b;
}

struct Range {
    int64_t lo = INT64_MAX, hi = INT64_MIN;  // [lo, hi) bytes
    void add(int64_t a, int64_t len) {
        lo = std::min(lo, a);
        hi = std::max(hi, a + len);
    }
};

// one wide launch: R row blocks x S splits
template <int JT, int W, int CG, int RA, int KS>
void walk_wide(const PassGeom& p, int64_t row_bytes, int64_t bq_bytes, int64_t part_words, const Variant& v,
               int64_t* stats) {
    using C = WideCfg<JT, W, CG, RA, KS>;
    const int64_t R = variant_rowblocks(true, v, p.nrows);
    const int S = v.S;
    const int64_t per = (p.nss + S - 1) / S;
    const int64_t grid = R * S;
    check(grid >= 1 && grid <= 0x7fffffffll, "grid");
    std::vector<uint8_t> seen((size_t)(R * S), 0);
    std::vector<int64_t> kcov((size_t)R, 0);
    Range rows, dig;
    constexpr int J16 = 32 * JT;
    // KS = 2: the lane offsets of both copy instructions of the first and the last
    // fragment (the chunk swizzle is not monotone in the lane; the offsets do not depend
    // on the block), walked once
    Range lane_off;
    if (KS == 2)
        for (int f : {0, C::kAFrags - 1})
            for (int i = 0; i < 2; ++i)
                for (int lane = 0; lane < 64; ++lane) lane_off.add(wide2_lane_off(f, i, lane, p.pitch), 16);
    check(KS == 1 || (lane_off.lo >= 0 && lane_off.hi - 16 < ((int64_t)1 << 32)), "lane offset outside 32 bits");
    for (int64_t b = 0; b < grid; ++b) {
        const WideBlock wb = wide_block(b, R, S, C::kRows, p.nss, per);
        check(wb.sp >= 0 && wb.sp < S && wb.rb >= 0 && wb.rb < R, "block coordinates out of range");
        check(!seen[(size_t)(wb.rb * S + wb.sp)]++, "two blocks share a (row block, split)");
        kcov[(size_t)wb.rb] += wb.nt;
        check(wb.nt % KS == 0 && wb.H0 % KS == 0, "split not in whole stages");
        const int64_t nt = wb.nt / KS;
        if (nt > 0 && KS == 1) {
            // copies at stages 0 .. nt - 1 (the ring's clamped copies repeat nt - 1)
            for (int64_t t : {(int64_t)0, nt - 1})
                for (int f : {0, C::kAFrags - 1})
                    for (int lane : {0, 63}) {
                        const int64_t base = wide_row_base(wb.blk_row, lane & 31, lane >> 5, wb.H0, p.pitch);
                        rows.add(base + wide_stage_step(t) + f * wide_frag_stride(p.pitch), 16);
                    }
        } else if (nt > 0) {
            const int64_t base = wide_row_base(wb.blk_row, 0, 0, wb.H0, p.pitch);
            for (int64_t t : {(int64_t)0, nt - 1}) {
                rows.add(base + wide2_stage_step(t) + lane_off.lo, 16);
                rows.add(base + wide2_stage_step(t) + lane_off.hi - 16, 16);
            }
        }
        if (nt > 0)
            for (int64_t t : {(int64_t)0, nt - 1})
                for (int chunk : {0, C::kBChunks - 1})
                    for (int lane : {0, 63}) dig.add(dig_stage(wb.H0 / KS + t, C::kB) + dig_lane(chunk, lane), 16);
        // stores: rows of the block below nrows, every digit column of the fragment
        for (int rg : {0, W / CG - 1})
            for (int a : {0, RA - 1})
                for (int vv : {0, 15})
                    for (int g : {0, 1}) {
                        const int64_t r = wide_out_row(wb.blk_row + rg * 32 * RA, a, vv, g);
                        if (r >= p.nrows) continue;
                        for (int col : {0, J16 - 1}) {
                            const int64_t i = part_idx(wb.sp, p.nrows, r, J16, col);
                            check(i >= 0 && i < part_words, "partial store outside the split-partial buffer");
                        }
                    }
    }
    for (int64_t r = 0; r < R; ++r) check(kcov[(size_t)r] == 2 * p.nss, "splits do not tile the K range");
    check(rows.lo >= 0 && rows.hi <= row_bytes, "packed-row read outside the padded layout");
    check(dig.lo >= 0 && dig.hi <= bq_bytes, "digit read outside the digit matrix");
    // the epilogue adds S slots of every row
    check(part_idx(S - 1, p.nrows, p.nrows - 1, J16, J16 - 1) < part_words, "epilogue read outside the partials");
    stats[0] += grid;
    stats[1] = std::max(stats[1], rows.hi);
    stats[2] = std::max(stats[2], dig.hi);
}

// one thin launch: stream-K grid G over R row blocks x n_st stages
template <int W, int RT, int NS, int SS>
void walk_thin(const PassGeom& p, int64_t row_bytes, int64_t bq_bytes, int64_t part_words, const Variant& v,
               int64_t* stats) {
    using C = ThinCfg<W, RT, SS>;
    check(p.nss % SS == 0, "thin shape needs whole stages");
    const int64_t R = variant_rowblocks(false, v, p.nrows);
    const int64_t n_st = p.nss / SS, U = R * n_st, G = v.G;
    const int P = v.S;
    check(G >= 1 && G <= U, "stream-K grid");
    check(P == thin_pieces(R, n_st, G), "pieces");
    // per row block: next slot expected, and whether its last piece zeroed the rest
    std::vector<int64_t> next_slot((size_t)R, 0);
    std::vector<uint8_t> closed((size_t)R, 0);
    std::vector<int64_t> stages((size_t)R, 0);
    Range rows, dig;
    const int64_t lane_lo = thin_lane_off(0, RT, 0, 0, p.pitch);
    const int64_t lane_hi = thin_lane_off(W - 1, RT, 15, 3, p.pitch);
    const int64_t frag_hi = thin_frag_off(RT - 1, SS - 1, p.pitch);
    for (int64_t b = 0; b < G; ++b) {
        const int64_t lo = sk_lo(b, U, G), hi = sk_lo(b + 1, U, G);
        const int64_t nt = hi - lo;
        if (nt <= 0) continue;
        // copies: NS - 1 ahead, then one per stage multiplied
        SkCursor cur(lo, nt, n_st);
        for (int64_t q = 0; q < nt + NS - 1; ++q) {
            check(cur.r < R, "copy past the last row block");
            const int64_t u = thin_unit_off(cur.r, C::kRB, cur.t, SS, p.pitch);
            rows.add(u + lane_lo, 16);
            rows.add(u + lane_hi + frag_hi, 16);
            dig.add(dig_stage(cur.t, C::kB) + dig_lane(0, 0), 16);
            dig.add(dig_stage(cur.t, C::kB) + dig_lane(C::kDigKB - 1, 63), 16);
            cur.advance();
        }
        // consumption and the stores at the end of each piece
        int64_t cr = lo / n_st, ct = lo % n_st;
        for (int64_t i = 0; i < nt; ++i) {
            ++stages[(size_t)cr];
            const bool row_end = ++ct == n_st;
            if (row_end || i == nt - 1) {
                const int64_t slot = sk_slot(b, cr, n_st, U, G);
                check(slot == next_slot[(size_t)cr] && !closed[(size_t)cr], "stream-K pieces out of order");
                check(slot >= 0 && slot < P, "stream-K slot outside [0, P)");
                ++next_slot[(size_t)cr];
                if (row_end) closed[(size_t)cr] = 1;
                for (int wave : {0, W - 1})
                    for (int a : {0, RT - 1})
                        for (int vv : {0, 3})
                            for (int g : {0, 3}) {
                                const int64_t r = thin_out_row(cr, C::kRB, wave, RT, a, vv, g);
                                if (r >= p.nrows) continue;
                                for (int64_t z : {slot, (int64_t)(row_end ? P - 1 : slot)})
                                    for (int rr : {0, 15}) {
                                        const int64_t idx = part_idx(z, p.nrows, r, 16, rr);
                                        check(idx >= 0 && idx < part_words, "partial store outside the buffer");
                                    }
                            }
                if (row_end) {
                    ct = 0;
                    ++cr;
                }
            }
        }
    }
    for (int64_t r = 0; r < R; ++r) {
        check(stages[(size_t)r] == n_st, "row block not covered by exactly its stages");
        check(closed[(size_t)r] && next_slot[(size_t)r] >= 1 && next_slot[(size_t)r] <= P,
              "row block's partial slots not all written");
    }
    check(rows.lo >= 0 && rows.hi <= row_bytes, "packed-row read outside the padded layout");
    check(dig.lo >= 0 && dig.hi <= bq_bytes, "digit read outside the digit matrix");
    check(part_idx(P - 1, p.nrows, p.nrows - 1, 16, 15) < part_words, "epilogue read outside the partials");
    stats[0] += G;
    stats[1] = std::max(stats[1], rows.hi);
    stats[2] = std::max(stats[2], dig.hi);
}

template <int JT, int Q>
void wide_q(const PassGeom& p, int64_t rb, int64_t bqb, int64_t pw, const Variant& v, int64_t* st) {
    if constexpr (wide_fits(JT, Q)) {
        constexpr WideShape w = kWide[Q];
        static_assert(w.NS * WideCfg<JT, w.W, w.CG, w.RA, w.KS>::kSlot + WideCfg<JT, w.W, w.CG, w.RA, w.KS>::kSink <=
                          160 * 1024,
                      "ring exceeds the LDS");
        walk_wide<JT, w.W, w.CG, w.RA, w.KS>(p, rb, bqb, pw, v, st);
    } else {
        check(false, "wide shape listed for a width it does not fit");
    }
}

template <int JT>
void wide_jt(const PassGeom& p, int64_t rb, int64_t bqb, int64_t pw, const Variant& v, int64_t* st) {
    switch (v.cfg) {
        case 0: wide_q<JT, 0>(p, rb, bqb, pw, v, st); break;
        case 1: wide_q<JT, 1>(p, rb, bqb, pw, v, st); break;
        case 2: wide_q<JT, 2>(p, rb, bqb, pw, v, st); break;
        case 3: wide_q<JT, 3>(p, rb, bqb, pw, v, st); break;
        case 4: wide_q<JT, 4>(p, rb, bqb, pw, v, st); break;
        case 5: wide_q<JT, 5>(p, rb, bqb, pw, v, st); break;
        case 6: wide_q<JT, 6>(p, rb, bqb, pw, v, st); break;
        case 7: wide_q<JT, 7>(p, rb, bqb, pw, v, st); break;
        case 8: wide_q<JT, 8>(p, rb, bqb, pw, v, st); break;
        case 9: wide_q<JT, 9>(p, rb, bqb, pw, v, st); break;
        case 10: wide_q<JT, 10>(p, rb, bqb, pw, v, st); break;
        case 11: wide_q<JT, 11>(p, rb, bqb, pw, v, st); break;
        case 12: wide_q<JT, 12>(p, rb, bqb, pw, v, st); break;
        default: wide_q<JT, 13>(p, rb, bqb, pw, v, st); break;
    }
}

template <int Q>
void thin_q(const PassGeom& p, int64_t rb, int64_t bqb, int64_t pw, const Variant& v, int64_t* st) {
    constexpr ThinShape t = kThin[Q];
    walk_thin<t.W, t.RT, t.NS, t.SS>(p, rb, bqb, pw, v, st);
}

void walk(bool wide, int JT, const PassGeom& p, int64_t rb, int64_t bqb, int64_t pw, const Variant& v,
          int64_t* st) {
    if (wide) {
        switch (JT) {
            case 1: wide_jt<1>(p, rb, bqb, pw, v, st); break;
            case 2: wide_jt<2>(p, rb, bqb, pw, v, st); break;
            case 3: wide_jt<3>(p, rb, bqb, pw, v, st); break;
            default: wide_jt<4>(p, rb, bqb, pw, v, st); break;
        }
        return;
    }
    switch (v.cfg) {
        case 0: thin_q<0>(p, rb, bqb, pw, v, st); break;
        case 1: thin_q<1>(p, rb, bqb, pw, v, st); break;
        case 2: thin_q<2>(p, rb, bqb, pw, v, st); break;
        case 3: thin_q<3>(p, rb, bqb, pw, v, st); break;
        case 4: thin_q<4>(p, rb, bqb, pw, v, st); break;
        case 5: thin_q<5>(p, rb, bqb, pw, v, st); break;
        case 6: thin_q<6>(p, rb, bqb, pw, v, st); break;
        default: thin_q<7>(p, rb, bqb, pw, v, st); break;
    }
}

// One gathered pass-1 launch (k_geno_tg<RB>) over ng listed individuals: R workgroups
// across the individual-major pitch, P splits of the stages. Walks every workgroup, copy
// lane and accumulator element through the kernel's index functions.
template <int RB>
void walk_tg(int64_t m, int64_t n, int64_t ng, int64_t pitch, int ncu, int64_t row_bytes, int64_t dig_bytes,
             int64_t part_bytes, int64_t* stats) {
    const TgLaunch tl = tg_launch<RB>(pitch, m, ng, ncu, part_bytes);
    const int64_t R = tl.R, P = tl.P, nst = ng / 64;
    constexpr int kTile = 64 * RB;
    check(R >= 1 && R * RB >= pitch && (R - 1) * RB < pitch, "workgroups do not tile the pitch");
    check(P >= 1 && P <= 8 && (P == 1 || nst / P >= 16), "split count");
    check(P * m * 16 * (int64_t)sizeof(int32_t) <= part_bytes, "partial slots exceed the partial buffer");
    // splits tile the stages
    for (int64_t sp = 0; sp < P; ++sp)
        check(tg_split_lo(sp + 1, nst, P) >= tg_split_lo(sp, nst, P), "splits out of order");
    check(tg_split_lo(0, nst, P) == 0 && tg_split_lo(P, nst, P) == nst, "splits do not tile the stages");
    // copies: a stage row is filled by one (copy, wave, lane) per LDS slot, and the slot
    // holds the global piece that the transposing reads of its logical piece expect
    std::vector<int> owner((size_t)kTile / 16, -1);
    for (int i = 0; i < tg_copies<RB>(); ++i)
        for (int w = 0; w < 8; ++w)
            for (int lane = 0; lane < 64; ++lane) {
                const int r = tg_copy_row<RB>(i, w, lane);
                check(r >= 0 && r < 64, "copy row outside the stage");
                const int64_t lds = 1024 * (8 * i + w) + 16 * lane;  // where the copy lands
                check(lds == (int64_t)r * RB + 16 * (lane % tg_pieces<RB>()), "copy lands off its row");
                int& o = owner[(size_t)(lds / 16)];
                check(o < 0, "two copies fill one LDS slot");
                o = (int)tg_piece_off<RB>(0, r, lane, INT64_MAX / 2);  // the piece's byte offset
            }
    for (int r = 0; r < 64; ++r)
        for (int pc = 0; pc < tg_pieces<RB>(); ++pc) {
            const int at = tg_row_slot<RB>(r, pc);
            check(at >= r * RB && at + 16 <= (r + 1) * RB, "read slot outside its row");
            check(owner[(size_t)(at / 16)] == 16 * pc, "read slot does not hold the piece it reads");
        }
    // transposing reads of a lane (group g, column sl): rows 16 g + sl / 2 and + 8
    for (int g = 0; g < 4; ++g)
        for (int sl = 0; sl < 16; ++sl) {
            const int rr = 16 * g + (sl >> 1);
            check(rr + 8 < 64, "transposing read past the stage");
            check(tg_dig_slot(sl, g) + 16 <= 1024, "digit read outside the stage's digits");
        }
    // digit copies: lanes 0..7 of every wave fill the 16 x 64 digit block once
    std::vector<int> dseen(16 * 4, 0);
    for (int w = 0; w < 8; ++w)
        for (int lane = 0; lane < 8; ++lane) {
            const int64_t off = tg_dig_off(w, lane, ng);
            const int nn = tg_dig_col(w, lane);
            check(nn >= 0 && nn < 16, "digit column");
            const int64_t chunk = (off - nn * ng) / 16;
            check(chunk >= 0 && chunk < 4 && (off - nn * ng) % 16 == 0, "digit chunk");
            ++dseen[(size_t)(nn * 4 + chunk)];
            // stage t of the split starts at individual 64 t
            check(off >= 0 && off + 64 * (nst - 1) + 16 <= 16 * ng && 16 * ng <= dig_bytes,
                  "digit read outside the digit matrix");
            // and lands where the reads of (column nn, chunk) look
            check(128 * w + 16 * lane == tg_dig_slot(nn, (int)chunk), "digit copy lands off its slot");
        }
    for (int v : dseen) check(v == 1, "digit block not filled exactly once");
    // row reads: every listed row < n, at a clamped piece of every workgroup
    Range rows;
    for (int64_t x = 0; x < R; ++x)
        for (int i = 0; i < tg_copies<RB>(); ++i)
            for (int w = 0; w < 8; ++w)
                for (int lane = 0; lane < 64; ++lane) {
                    const int r = tg_copy_row<RB>(i, w, lane);
                    const int64_t off = tg_piece_off<RB>(x * RB, r, lane, pitch);
                    check(off >= 0 && off + 16 <= pitch, "piece outside its row");
                    rows.add(off, 16);
                    rows.add((n - 1) * pitch + off, 16);
                }
    check(rows.lo >= 0 && rows.hi <= row_bytes, "packed-row read outside the padded layout");
    // index reads: stage rows of stages 0 .. nst - 1 (later stages clamp to the last)
    check(64 * (nst - 1) + 63 < ng, "index read past the listed individuals");
    // stores: every SNP < m of every split exactly once per column
    std::vector<uint8_t> hit((size_t)(P * m), 0);
    for (int64_t sp = 0; sp < P; ++sp)
        for (int64_t x = 0; x < R; ++x)
            for (int w = 0; w < 8; ++w)
                for (int i = 0; i < tg_copies<RB>(); ++i)
                    for (int g = 0; g < 4; ++g)
                        for (int v = 0; v < 4; ++v)
                            for (int j = 0; j < 4; ++j) {
                                const int64_t snp = tg_out_snp<RB>(x * RB, w, i, g, v, j);
                                check(snp >= 0 && snp < 4 * R * RB, "SNP outside the workgroups");
                                if (snp >= m) continue;
                                for (int sl : {0, 15}) {
                                    const int64_t idx = part_idx(sp, m, snp, 16, sl);
                                    check(idx >= 0 && (idx + 1) * 4 <= part_bytes, "partial store outside the buffer");
                                }
                                check(!hit[(size_t)(sp * m + snp)]++, "SNP stored twice by a split");
                            }
    for (uint8_t h : hit) check(h == 1, "SNP not stored by a split");
    check((part_idx(P - 1, m, m - 1, 16, 15) + 1) * 4 <= part_bytes, "epilogue read outside the partials");
    stats[0] = R;
    stats[1] = P;
    stats[2] = rows.hi - row_bytes;
}

}  // namespace

extern "C" {

// Replays the gathered pass 1 of a thin product (k_quantize_gather + k_geno_tg<256>) over
// ng listed individuals (a positive multiple of 512, at most n + 512) of a cohort of n
// individuals and m kept SNPs, with the buffers mfma_geno_product allocates for it
// (the digit matrix and split partials of the product's geometry on ng rows; the launch
// takes fewer splits from a smaller partial buffer). stats (5): [0] workgroups across the
// pitch, [1] splits, [2] largest packed-row byte read + 1 over the layout's bytes, [3]
// digit bytes, [4] partial bytes. shrink (3, may be NULL) as for globe_plan_replay.
// Returns 0, or 1 with the first violation in msg.
int globe_plan_replay_tg(int64_t m, int64_t n, int64_t ng, int ncu, const int* thin_bpc, const int64_t* shrink,
                         int64_t* stats, char* msg, int msglen) {
    std::string err;
    std::memset(stats, 0, 5 * sizeof(int64_t));
    try {
        check(m >= 1 && n >= 1 && ng > 0 && ng % kRowPad == 0 && ng <= n + kRowPad && ncu >= 1, "arguments");
        const int64_t pitch_s = packed_pitch(n), pitch_i = packed_pitch(m);
        const Geometry geo = contract_geometry(16, m, ng, pitch_s, pitch_i, 0, ncu, thin_bpc);
        check(!geo.wide && geo.J16 == 16, "a gathered pass 1 is thin");
        const int64_t sh[3] = {shrink ? shrink[0] : 0, shrink ? shrink[1] : 0, shrink ? shrink[2] : 0};
        const int64_t row_bytes = packed_rows(n) * pitch_i - sh[0];
        const int64_t dig_bytes = geo.bq_bytes - sh[1];
        // the digit transpose: column nn of listed individual t at nn ng + t
        check(15 * ng + ng - 1 < dig_bytes, "gathered digit write outside the digit matrix");
        stats[3] = geo.bq_bytes;
        stats[4] = geo.part_bytes;
        walk_tg<256>(m, n, ng, pitch_i, ncu, row_bytes, dig_bytes, geo.part_bytes - sh[2], stats);
    } catch (const Fail& f) {
        err = f.msg;
    }
    if (msg && msglen > 0) std::snprintf(msg, (size_t)msglen, "%s", err.c_str());
    return err.empty() ? 0 : 1;
}

// Replays both passes of one product shape: J digit columns (1..128), m kept SNPs, n
// individuals, pass 2 from row r0 (the GRM builder; 0 otherwise), ncu compute units and
// thin_bpc[8] resident thin blocks per CU. stats (8): [0] candidates of pass 1, [1] of
// pass 2, [2] blocks walked, [3] largest packed-row byte read + 1 over the layout's
// bytes (pass 1), [4] the same for pass 2, [5] digit bytes, [6] partial bytes, [7]
// wide (1) or thin (0). shrink (3, may be NULL): bytes taken off the packed layouts, the
// digit matrix and the partial buffer, so a test can show that an allocation one access
// too small is caught. Returns 0, or 1 with the first violation in msg.
int globe_plan_replay(int J, int64_t m, int64_t n, int64_t r0, int ncu, const int* thin_bpc, const int64_t* shrink,
                      int64_t* stats, char* msg, int msglen) {
    std::string err;
    std::memset(stats, 0, 8 * sizeof(int64_t));
    try {
        check(J >= 1 && J <= 128 && m >= 1 && n >= 1 && r0 >= 0 && r0 < n && r0 % kRowPad == 0, "arguments");
        const int64_t pitch_s = packed_pitch(n), pitch_i = packed_pitch(m);
        const Geometry geo = contract_geometry(J, m, n, pitch_s, pitch_i, r0, ncu, thin_bpc);
        // bytes readable from each pass's row pointer: the padded layouts, pass 2 from row r0
        const int64_t sh[3] = {shrink ? shrink[0] : 0, shrink ? shrink[1] : 0, shrink ? shrink[2] : 0};
        const int64_t row_bytes[2] = {packed_rows(m) * pitch_s - sh[0], (packed_rows(n) - r0) * pitch_i - sh[0]};
        const int64_t bq_bytes = geo.bq_bytes - sh[1];
        const int64_t part_words = (geo.part_bytes - sh[2]) / 4;
        // quantize: every padded genotype of every digit column lands inside Bq
        for (int q = 0; q < 2; ++q) {
            const int fw = geo.wide ? 32 : 16;
            for (int64_t kap : {(int64_t)0, geo.pass[q].Kpad - 1})
                for (int j : {0, geo.J16 - 1}) {
                    const int64_t i = bq_index(kap, j % fw, j / fw, geo.nfrag, geo.wide);
                    check(i >= 0 && i < bq_bytes, "quantize write outside the digit matrix");
                }
        }
        stats[3] = stats[4] = INT64_MIN;
        for (int q = 0; q < 2; ++q) {
            const std::vector<Variant> cand = candidates(geo.wide, geo.JT, geo.pass[q].nrows, geo.pass[q].nss, ncu,
                                                         thin_bpc);
            check(!cand.empty(), "no candidate");
            stats[q] = (int64_t)cand.size();
            for (const Variant& v : cand) {
                int64_t st[3] = {0, 0, 0};
                try {
                    walk(geo.wide, geo.JT, geo.pass[q], row_bytes[q], bq_bytes, part_words, v, st);
                } catch (const Fail& f) {
                    throw Fail{vname(q, geo.wide, v) + ": " + f.msg};
                }
                stats[2] += st[0];
                stats[3 + q] = std::max(stats[3 + q], st[1] - row_bytes[q]);
                stats[5] = geo.bq_bytes;
            }
        }
        stats[6] = geo.part_bytes;
        stats[7] = geo.wide ? 1 : 0;
    } catch (const Fail& f) {
        err = f.msg;
    }
    if (msg && msglen > 0) {
        std::snprintf(msg, (size_t)msglen, "%s", err.c_str());
    }
    return err.empty() ? 0 : 1;
}

}  // extern "C"
