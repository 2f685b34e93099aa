// Dense symmetric GRM product GZ = sig2g G S0 with FP64 accumulation, and the GRM
// builder. The reference keeps G as the FP32 packed lower triangle of a GCTA .grm.bin
// (element (i, j), j <= i, at i(i+1)/2 + j; PCGsolver.cpp:878-881, :1204-1212;
// varCompEst.cpp:336-347). On the device the same FP32 values are kept tile-major:
// 64 x 64 tiles (I, J), J <= I, tile index I(I+1)/2 + J, each tile row-major and 16 KB
// contiguous; the diagonal tiles hold both halves (mirrored), and rows or columns past n
// are zero. A packed row of the triangle starts at i(i+1)/2 floats, which is not 16-byte
// aligned, so tiles could only be gathered 4 bytes per lane; tile-major, every tile
// streams as 1 KB pieces of 16 bytes per lane. Conversion to and from the packed triangle
// happens at load (globe_load_grm) and read-out (globe_get_grm).
//
// Product: a workgroup owns a strip of 64 output rows and one of kGrmChunks ranges of the
// strip's tiles J: tile (I, J) for J <= I, tile (J, I) read transposed for J > I. Tiles
// and the matching 64-row slice of S0 (pre-transposed to rows of 16, 20, 24 or 32
// doubles) go global -> LDS (global_load_lds_dwordx4) into a ring of kGS slots, one wait +
// barrier per tile, and each tile is multiplied on the FP64 matrix cores while the next
// one lands. A group of up to 32 columns takes one pass over the triangle: 16 columns on
// the 16 x 16 MFMA; 17-24 with one or two more groups of 4 on the 4 x 4 MFMA; 25-32 in
// blocks of 8 waves (two column groups of 16 share each tile). The strip partials of the
// chunks are added in chunk order, so the result is deterministic. Each stored tile is
// read twice (once for its row strip, once for its column strip); the triangle is far
// larger than any cache. Products of at most 8 columns read each stored tile once
// instead: its transposed product goes to a per-tile partial of 64 x k doubles that a
// finish kernel adds per strip in fixed order (k_grm_sym, k_grm_symx).
#include <algorithm>

#include "globe_internal.h"

namespace globe {

namespace {

constexpr int kGT = 64;        // tile edge
constexpr int kGrmChunks = 4;  // column ranges per row strip
constexpr int kGC = 16;        // columns of S0 per column group (one MFMA column block)
constexpr int kGCMax = 32;     // columns per pass: up to two column groups
constexpr int kGS = 2;         // ring slots: one tile in flight while one is multiplied
constexpr int kTileB = kGT * kGT * 4;                 // FP32 tile bytes

// A block of 4 CG waves: wave w owns rows 16 (w % 4) .. +15 of the strip and column
// group w / 4 (16 columns); the CG groups share each tile. E > 0 (with CG = 1): each wave
// also owns E groups of 4 more columns (16 + 4 e ..) on the 4 x 4 FP64 MFMA, so 17-24
// columns cost 1.25-1.5 times the MFMA work of 16 instead of twice.
template <int CG, int E = 0>
struct GrmCfg {
    static_assert(E == 0 || CG == 1, "4-column groups extend one 16-column group");
    static constexpr int kCols = kGC * CG + 4 * E;        // columns per pass (S0 slice row)
    static constexpr int kWaves = 4 * CG;
    static constexpr int kSliceB = kGT * kCols * 8;       // S0 slice bytes
    static constexpr int kSlotB = kTileB + kSliceB;
    static constexpr int kTilePerWave = 16 / kWaves;      // 1 KB tile pieces per wave
    static constexpr int kSlicePieces = kSliceB / 1024;   // 1 KB slice pieces, strided over the waves
    static constexpr int kSlicePerWave = (kSlicePieces + kWaves - 1) / kWaves;
    static constexpr int kCopies = kTilePerWave + kSlicePerWave;
    static_assert(16 % kWaves == 0 && kSliceB % 1024 == 0, "whole 1 KB pieces");
    // with more than two ring slots the wait counts a wave's copies of later stages, which
    // needs every wave to issue the same number
    static_assert(kGS == 2 || kSlicePieces % kWaves == 0, "pieces spread evenly");
};

typedef double v4d __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

__device__ inline int64_t lt_index(int64_t i, int64_t j) { return i * (i + 1) / 2 + j; }
__device__ inline int64_t tile_index(int64_t I, int64_t J) { return I * (I + 1) / 2 + J; }  // J <= I

// Element (a, b) of a tile in LDS: row a holds its 16 pieces of 4 floats rotated by a,
// piece q at slot (q + a) & 15. 16 consecutive rows at one column (the MFMA A operand)
// and 16 consecutive columns of one row (the transposed read) fall in different banks.
__device__ inline int rot(int a, int b) { return a * kGT + ((((b >> 2) + a) & 15) << 2) + (b & 3); }

// S0T[j][c] = S0[c][j] for c < kk and j < n, 0 elsewhere (npad = 64 ntiles rows of
// `cols` doubles).
__global__ __launch_bounds__(256) void k_grm_s0t(const double* __restrict__ S0, int64_t ldn, int32_t n,
                                                 int kk, int cols, int64_t npad, double* __restrict__ S0T,
                                                 const int* skip) {
    if (skip && *skip) return;
    for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < npad * cols; e += (int64_t)gridDim.x * 256) {
        const int64_t j = e / cols;
        const int c = (int)(e % cols);
        S0T[e] = (c < kk && j < n) ? S0[c * ldn + j] : 0.0;
    }
}

// Copies of one stage into ring slot `slot` by wave w: tile pieces (1 KB = 4 tile rows,
// lane l: row 4 p + l / 16, its piece (l % 16 - row) & 15, so the rows land rotated)
// p = kTilePerWave w .., and the 1 KB pieces w, w + kWaves, .. of the S0T slice.
template <int CG, int E>
__device__ inline void grm_issue(const float* __restrict__ tile, const double* __restrict__ s0t,
                                 uint8_t* __restrict__ slot, int w, int lane) {
    using C = GrmCfg<CG, E>;
#pragma unroll
    for (int q = 0; q < C::kTilePerWave; ++q) {
        const int p = C::kTilePerWave * w + q;
        const int a = 4 * p + (lane >> 4);
        const int piece = ((lane & 15) - a) & 15;
        __builtin_amdgcn_global_load_lds((const void*)(tile + a * kGT + 4 * piece),
                                         (lds_void*)(slot + p * 1024), 16, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < C::kSlicePerWave; ++q) {
        const int p = w + C::kWaves * q;
        if (p < C::kSlicePieces)  // uniform in the wave
            __builtin_amdgcn_global_load_lds((const void*)((const uint8_t*)s0t + p * 1024 + lane * 16),
                                             (lds_void*)(slot + kTileB + p * 1024), 16, 0, 0);
    }
}

// 16 MFMAs of one tile for the wave of row group rg and column group cg: out rows
// 16 rg + (0..15) x columns 16 cg + (0..15). A lane l = G[16 rg + l % 16][4 s + l / 16]
// (TRANS: the stored tile read transposed), B lane l = S0[j0 + 4 s + l / 16][16 cg + l % 16],
// D element v of lane l = (row 4 v + l / 16, column l % 16) (v_mfma_f64_16x16x4f64,
// probed on gfx950).
// With E > 0, 16 E more of the 4-block v_mfma_f64_4x4x4f64 (probed on gfx950: block b of
// lanes 4 b .. 4 b + 3 + 16 kk; A lane 16 kk + 4 b + r = (row r, k kk) of block b, B lane
// 16 kk + 4 b + c = (k kk, column c), D lane 16 r + 4 b + c = (row r, column c)): block b
// takes rows 16 rg + 4 b .. + 3, so its A operand is the 16 x 16 form's, lane for lane, and
// every block reads the same 4 columns 16 + 4 e .. of S0; accx[e] of lane l = (row
// 16 rg + 4 ((l >> 2) & 3) + l / 16, column 16 + 4 e + l % 4).
template <int CG, int E, bool TRANS>
__device__ inline void grm_mma(const uint8_t* __restrict__ slot, int rg, int cg, int lane, v4d& acc,
                               double (&accx)[2]) {
    constexpr int kCols = GrmCfg<CG, E>::kCols;
    const float* g = reinterpret_cast<const float*>(slot);
    const double* srow = reinterpret_cast<const double*>(slot + kTileB) + (lane >> 4) * kCols;
    const double* sb = srow + kGC * cg + (lane & 15);
    const double* sx = srow + kGC + (lane & 3);
    const int a = 16 * rg + (lane & 15);
#pragma unroll
    for (int s4 = 0; s4 < kGT / 4; ++s4) {
        const int b = 4 * s4 + (lane >> 4);
        const double gv = (double)(TRANS ? g[rot(b, a)] : g[rot(a, b)]);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(gv, sb[4 * s4 * kCols], acc, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < E; ++e)
            accx[e] = __builtin_amdgcn_mfma_f64_4x4x4f64(gv, sx[4 * s4 * kCols + 4 * e], accx[e], 0, 0, 0);
    }
}

// One pass of up to 16 CG + 4 E columns (S0T zero past kk): part[h][c][i] = sum over the
// tiles J of chunk h of (G_IJ S0_J)[i][c].
template <int CG, int E>
__global__ __launch_bounds__(256 * CG) void k_grm_tiles(const float* __restrict__ tiles, int32_t n,
                                                        const double* __restrict__ S0T, int kk,
                                                        int32_t ntiles, double* __restrict__ part,
                                                        const int* skip) {
    if (skip && *skip) return;
    using C = GrmCfg<CG, E>;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[kGS * C::kSlotB];
    const int I = blockIdx.x, h = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rg = w & 3, cg = w >> 2;
    const int J0 = (int)((int64_t)h * ntiles / kGrmChunks), J1 = (int)((int64_t)(h + 1) * ntiles / kGrmChunks);
    const int nt = J1 - J0;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    double accx[2] = {0.0, 0.0};
    if (nt > 0) {
        // stage t: tile J0 + t (repeating the last one past the end, into slots never read)
        auto issue = [&](int t, int slot) {
            const int64_t J = J0 + (t < nt ? t : nt - 1);
            const float* tile = tiles + (J <= I ? tile_index(I, J) : tile_index(J, I)) * (kGT * kGT);
            grm_issue<CG, E>(tile, S0T + J * kGT * C::kCols, lds + slot * C::kSlotB, w, lane);
        };
#pragma unroll
        for (int u = 0; u < kGS - 1; ++u) issue(u, u);
        int cs = 0;
        for (int t = 0; t < nt; ++t) {
            // this wave's copies of stage t have landed (those of the kGS - 2 later stages
            // may be in flight); after the barrier every wave's have, and every wave is done
            // with the slot the next copies overwrite
            asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"((kGS - 2) * C::kCopies)
                         : "memory");
            issue(t + kGS - 1, cs == 0 ? kGS - 1 : cs - 1);
            __builtin_amdgcn_sched_barrier(0);
            const uint8_t* slot = lds + cs * C::kSlotB;
            if (J0 + t <= I) grm_mma<CG, E, false>(slot, rg, cg, lane, acc, accx);
            else grm_mma<CG, E, true>(slot, rg, cg, lane, acc, accx);
            cs = cs + 1 == kGS ? 0 : cs + 1;
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    const int col = kGC * cg + (lane & 15);
    const int64_t i0 = (int64_t)I * kGT;
    if (col < kk) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t i = i0 + 16 * rg + 4 * v + (lane >> 4);
            if (i < n) part[((int64_t)h * kGCMax + col) * n + i] = acc[v];
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int cx = kGC + 4 * e + (lane & 3);
        const int64_t i = i0 + 16 * rg + 4 * ((lane >> 2) & 3) + (lane >> 4);
        if (cx < kk && i < n) part[((int64_t)h * kGCMax + cx) * n + i] = accx[e];
    }
}

// GZ[c][i] = sig2g_c * sum over chunks of the strip partials (fixed order)
__global__ __launch_bounds__(256) void k_grm_finish(const double* __restrict__ part, int32_t n, int kc,
                                                    const double* __restrict__ th,
                                                    double* __restrict__ GZ, int64_t ldg,
                                                    const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double sg = th[c * 5 + 1];
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double v = 0.0;
#pragma unroll
        for (int h = 0; h < kGrmChunks; ++h) v += part[((int64_t)h * kc + c) * n + i];
        GZ[c * ldg + i] = sg * v;
    }
}

// One- or two-column product that reads every stored tile once. A workgroup owns strip I
// and one of kSymChunks ranges of J <= I (the longest strips are dispatched first); wave
// w holds rows 16 w .. 16 w + 15 of each tile, lane = the tile column, so a row of a
// tile is one 256-byte wave load. Per tile (I, J):
//   rows:    acc[r][q] += G_IJ[r][lane] x_J[lane][q], summed over the lanes once per strip
//            into rowp[h][q][i] (every workgroup writes its rows, zeros for an empty range);
//   columns: for J < I, sum_r G_IJ[r][lane] x_I[r][q] over the 4 waves (LDS, fixed order)
//            into colp[tile (I, J)][q][lane], the strip-J partial of the transposed tile.
// The diagonal tile is stored whole and only contributes to rows. The finish adds the
// chunk partials and the column partials of strip J in fixed order: deterministic.
constexpr int kSymChunks = 8;

template <int KC>
__global__ __launch_bounds__(256) void k_grm_sym(const float* __restrict__ tiles, int32_t n, int32_t ntiles,
                                                 const double* __restrict__ S0, int64_t ldn, int kk,
                                                 double* __restrict__ rowp, double* __restrict__ colp,
                                                 const int* skip) {
    if (skip && *skip) return;
    const int32_t I = ntiles - 1 - (int32_t)blockIdx.x;
    const int h = blockIdx.y;
    const int32_t J0 = (int32_t)((int64_t)h * (I + 1) / kSymChunks);
    const int32_t J1 = (int32_t)((int64_t)(h + 1) * (I + 1) / kSymChunks);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ double xi[kGT][KC];
    __shared__ double cred[2][4][kGT][KC];
    for (int e = threadIdx.x; e < kGT * KC; e += 256) {
        const int r = e / KC, q = e % KC;
        const int64_t i = (int64_t)I * kGT + r;
        xi[r][q] = (q < kk && i < n) ? S0[q * ldn + i] : 0.0;
    }
    __syncthreads();
    double acc[16][KC];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr)
#pragma unroll
        for (int q = 0; q < KC; ++q) acc[rr][q] = 0.0;
    for (int32_t J = J0; J < J1; ++J) {
        const float* t = tiles + tile_index(I, J) * (kGT * kGT) + (16 * w) * kGT + lane;
        float gv[16];
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) gv[rr] = t[rr * kGT];
        const int64_t j = (int64_t)J * kGT + lane;
        double xj[KC], cp[KC];
#pragma unroll
        for (int q = 0; q < KC; ++q) {
            xj[q] = (q < kk && j < n) ? S0[q * ldn + j] : 0.0;
            cp[q] = 0.0;
        }
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            const double g = (double)gv[rr];
#pragma unroll
            for (int q = 0; q < KC; ++q) {
                acc[rr][q] += g * xj[q];
                cp[q] += g * xi[16 * w + rr][q];
            }
        }
        if (J < I) {
            // double-buffered: the reduction of tile J reads cred[J & 1] while tile J + 1
            // fills the other half; the barrier of J + 1 orders it before J + 2 refills
            const int b = J & 1;
#pragma unroll
            for (int q = 0; q < KC; ++q) cred[b][w][lane][q] = cp[q];
            __syncthreads();
            if (threadIdx.x < kGT) {
                double* dst = colp + tile_index(I, J) * (kGT * KC);
#pragma unroll
                for (int q = 0; q < KC; ++q)
                    dst[q * kGT + lane] = ((cred[b][0][lane][q] + cred[b][1][lane][q]) + cred[b][2][lane][q]) +
                                          cred[b][3][lane][q];
            }
        }
    }
    // each of the wave's 16 rows: the sum over the 64 lanes
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
#pragma unroll
        for (int q = 0; q < KC; ++q) {
            double v = acc[rr][q];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
            const int64_t i = (int64_t)I * kGT + 16 * w + rr;
            if (lane == 0 && q < kk && i < n) rowp[((int64_t)h * KC + q) * n + i] = v;
        }
    }
}

// GZ[q][i] = sig2g_q (sum_h rowp[h][q][i] + sum_{I > J} colp[tile (I, J)][q][a]) for the
// rows i = 64 J + a of strip J = blockIdx.x; wave w adds I = J + 1 + w, J + 5 + w, ...
// (kc: the partials' column count)
__global__ __launch_bounds__(256) void k_grm_sym_finish(const double* __restrict__ rowp,
                                                        const double* __restrict__ colp, int32_t n,
                                                        int32_t ntiles, int kc, const double* __restrict__ th,
                                                        double* __restrict__ GZ, int64_t ldg, const int* skip) {
    if (skip && *skip) return;
    const int32_t J = blockIdx.x;
    const int q = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ double red[4][kGT];
    double v = 0.0;
    for (int32_t I = J + 1 + w; I < ntiles; I += 4) v += colp[tile_index(I, J) * (kGT * kc) + q * kGT + lane];
    red[w][lane] = v;
    __syncthreads();
    if (w != 0) return;
    const int64_t i = (int64_t)J * kGT + lane;
    if (i >= n) return;
    double r = 0.0;
#pragma unroll
    for (int h = 0; h < kSymChunks; ++h) r += rowp[((int64_t)h * kc + q) * n + i];
    const double c = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    GZ[q * ldg + i] = th[q * 5 + 1] * (r + c);
}

// Few-column product (3 <= k <= 8) on the 4 x 4 FP64 MFMA, reading every stored tile once.
// Work split as in k_grm_sym (strip I, one of kSymChunks ranges of J <= I, the longest
// strips first); the tiles and the matching 64-row slices of S0T (rows of 4 E doubles)
// stream global -> LDS through a 2-slot ring as in k_grm_tiles, and x_I (the strip's own
// slice) sits in LDS. Per tile (I, J), wave w:
//   rows:    rows 16 w .. + 15 of G_IJ x_J (16 E MFMAs over the tile's 64 columns), kept in
//            registers over the range and written once into rowp[h][q][i];
//   columns: for J < I, tile columns 16 w .. + 15 of G_IJ^T x_I (16 E MFMAs over the
//            tile's 64 rows, the transposed LDS read), complete for the tile, written into
//            colp[tile (I, J)][q][a].
// Lane layout of v_mfma_f64_4x4x4f64 as in grm_mma: block b of an instruction takes rows
// (or tile columns) 16 w + 4 b .. + 3. The partials have k_grm_sym's layout with kk
// columns, so k_grm_sym_finish adds them in the same fixed order.
template <int KC>  // 4 or 8: columns of the MFMA groups (kk <= KC)
__global__ __launch_bounds__(256) void k_grm_symx(const float* __restrict__ tiles, int32_t n, int32_t ntiles,
                                                  const double* __restrict__ S0T, int kk,
                                                  double* __restrict__ rowp, double* __restrict__ colp,
                                                  const int* skip) {
    if (skip && *skip) return;
    constexpr int E = (KC + 3) / 4;       // 4-column MFMA groups
    constexpr int kC4 = 4 * E;             // S0T row length
    constexpr int kSliceB = kGT * kC4 * 8;
    constexpr int kSlotB = kTileB + kSliceB;
    static_assert(kSliceB % 1024 == 0 && kSliceB / 1024 <= 4, "one slice piece per wave at most");
    __shared__ __attribute__((aligned(1024))) uint8_t lds[kGS * kSlotB + kSliceB];
    const int32_t I = ntiles - 1 - (int32_t)blockIdx.x;
    const int h = blockIdx.y;
    const int32_t J0 = (int32_t)((int64_t)h * (I + 1) / kSymChunks);
    const int32_t J1 = (int32_t)((int64_t)(h + 1) * (I + 1) / kSymChunks);
    const int nt = J1 - J0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double accr[E];
#pragma unroll
    for (int e = 0; e < E; ++e) accr[e] = 0.0;
    if (nt > 0) {
        auto issue = [&](int t, int slot) {
            const int64_t J = J0 + (t < nt ? t : nt - 1);
            const float* tile = tiles + tile_index(I, J) * (kGT * kGT);
            uint8_t* dst = lds + slot * kSlotB;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int p = 4 * w + q;
                const int a = 4 * p + (lane >> 4);
                const int piece = ((lane & 15) - a) & 15;
                __builtin_amdgcn_global_load_lds((const void*)(tile + a * kGT + 4 * piece),
                                                 (lds_void*)(dst + p * 1024), 16, 0, 0);
            }
            if (w < kSliceB / 1024)
                __builtin_amdgcn_global_load_lds(
                    (const void*)((const uint8_t*)(S0T + J * kGT * kC4) + w * 1024 + lane * 16),
                    (lds_void*)(dst + kTileB + w * 1024), 16, 0, 0);
        };
        issue(0, 0);
        // x_I: the strip's own slice, read by the column products
        double* xi = reinterpret_cast<double*>(lds + kGS * kSlotB);
        for (int e = threadIdx.x; e < kGT * kC4; e += 256) xi[e] = S0T[(int64_t)I * kGT * kC4 + e];
        const double* xis = xi + (lane >> 4) * kC4 + (lane & 3);
        int cs = 0;
        for (int t = 0; t < nt; ++t) {
            // this wave's copies of stage t have landed; after the barrier every wave's
            // have (and x_I is in place), and every wave is done with the other slot
            asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
            issue(t + 1, cs ^ 1);
            __builtin_amdgcn_sched_barrier(0);
            const uint8_t* slot = lds + cs * kSlotB;
            const float* g = reinterpret_cast<const float*>(slot);
            const double* xj = reinterpret_cast<const double*>(slot + kTileB) + (lane >> 4) * kC4 + (lane & 3);
            const int a = 16 * w + (lane & 15);
#pragma unroll
            for (int s4 = 0; s4 < kGT / 4; ++s4) {
                const double gv = (double)g[rot(a, 4 * s4 + (lane >> 4))];
#pragma unroll
                for (int e = 0; e < E; ++e)
                    accr[e] = __builtin_amdgcn_mfma_f64_4x4x4f64(gv, xj[4 * s4 * kC4 + 4 * e], accr[e], 0, 0, 0);
            }
            const int32_t J = J0 + t;
            if (J < I) {
                double accc[E];
#pragma unroll
                for (int e = 0; e < E; ++e) accc[e] = 0.0;
#pragma unroll
                for (int s4 = 0; s4 < kGT / 4; ++s4) {
                    const double gv = (double)g[rot(4 * s4 + (lane >> 4), a)];
#pragma unroll
                    for (int e = 0; e < E; ++e)
                        accc[e] = __builtin_amdgcn_mfma_f64_4x4x4f64(gv, xis[4 * s4 * kC4 + 4 * e], accc[e], 0, 0, 0);
                }
                double* dst = colp + tile_index(I, J) * (kGT * kk);
                const int col = 16 * w + 4 * ((lane >> 2) & 3) + (lane >> 4);
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int q = 4 * e + (lane & 3);
                    if (q < kk) dst[q * kGT + col] = accc[e];
                }
            }
            cs ^= 1;
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    const int64_t i = (int64_t)I * kGT + 16 * w + 4 * ((lane >> 2) & 3) + (lane >> 4);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int q = 4 * e + (lane & 3);
        if (q < kk && i < n) rowp[((int64_t)h * kk + q) * n + i] = accr[e];
    }
}

// Packed rows r0 .. r1 - 1 of the triangle (lt holds them from lt_index(r0, 0) on) into
// the tiles; one block per row. The diagonal tile gets both (a, b) and (b, a).
__global__ __launch_bounds__(256) void k_grm_tiles_from_lt(const float* __restrict__ lt, int64_t r0,
                                                           float* __restrict__ tiles) {
    const int64_t i = r0 + blockIdx.x;
    const float* row = lt + (lt_index(i, 0) - lt_index(r0, 0));
    const int64_t I = i / kGT;
    const int a = (int)(i % kGT);
    for (int64_t j = threadIdx.x; j <= i; j += 256) {
        const int64_t J = j / kGT;
        const int b = (int)(j % kGT);
        float* t = tiles + tile_index(I, J) * (kGT * kGT);
        const float v = row[j];
        t[a * kGT + b] = v;
        if (J == I) t[b * kGT + a] = v;
    }
}

// The inverse: packed rows r0 .. r1 - 1 from the tiles into lt (from lt_index(r0, 0) on).
__global__ __launch_bounds__(256) void k_grm_lt_from_tiles(const float* __restrict__ tiles, int64_t r0,
                                                           float* __restrict__ lt) {
    const int64_t i = r0 + blockIdx.x;
    float* row = lt + (lt_index(i, 0) - lt_index(r0, 0));
    const int64_t I = i / kGT;
    const int a = (int)(i % kGT);
    for (int64_t j = threadIdx.x; j <= i; j += 256)
        row[j] = tiles[tile_index(I, j / kGT) * (kGT * kGT) + a * kGT + (j % kGT)];
}

// Columns j0 .. j0 + k - 1 of G (Gc: n x k, column-major; j0 a multiple of 64, k <= 64)
// into tile column J = j0 / 64, rounded to FP32: rows i >= j, mirrored in the diagonal
// tile.
__global__ __launch_bounds__(256) void k_grm_scatter(const double* __restrict__ Gc, int64_t ldg,
                                                     int32_t n, int64_t j0, int k,
                                                     float* __restrict__ tiles) {
    const int c = blockIdx.y;
    const int64_t j = j0 + c, J = j0 / kGT;
    for (int64_t i = j + blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t I = i / kGT;
        const int a = (int)(i % kGT);
        float* t = tiles + tile_index(I, J) * (kGT * kGT);
        const float v = (float)Gc[c * ldg + i];
        t[a * kGT + c] = v;
        if (I == J) t[c * kGT + a] = v;
    }
}

__global__ __launch_bounds__(256) void k_grm_diag(const float* __restrict__ tiles, int32_t n,
                                                  double* __restrict__ d) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t I = i / kGT;
        const int a = (int)(i % kGT);
        d[i] = (double)tiles[tile_index(I, I) * (kGT * kGT) + a * kGT + a];
    }
}

}  // namespace

size_t grm_tile_bytes(int32_t n) {
    const size_t T = (size_t)(n + kGT - 1) / kGT;
    return T * (T + 1) / 2 * kTileB;
}

void launch_grm_tiles_from_lt(hipStream_t s, const float* lt_rows, int64_t r0, int64_t r1, float* tiles) {
    if (r1 <= r0) return;
    k_grm_tiles_from_lt<<<(unsigned)(r1 - r0), 256, 0, s>>>(lt_rows, r0, tiles);
    HIP_CHECK(hipGetLastError());
}

void launch_grm_lt_from_tiles(hipStream_t s, const float* tiles, int64_t r0, int64_t r1, float* lt_rows) {
    if (r1 <= r0) return;
    k_grm_lt_from_tiles<<<(unsigned)(r1 - r0), 256, 0, s>>>(tiles, r0, lt_rows);
    HIP_CHECK(hipGetLastError());
}

void launch_grm_diag(hipStream_t s, const float* tiles, int32_t n, double* d) {
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024));
    k_grm_diag<<<gx, 256, 0, s>>>(tiles, n, d);
    HIP_CHECK(hipGetLastError());
}

void launch_grm_product(hipStream_t s, const float* tiles, int32_t n, const double* S0, int64_t ldn,
                        int k, const double* theta_cols, double* GZ, int64_t ldg, const int* skip,
                        double* part) {
    if (n == 0) return;
    const int32_t ntiles = (n + kGT - 1) / kGT;
    const dim3 g(ntiles, kGrmChunks);
    const unsigned gf = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    // one pass over the triangle per group of up to 32 columns (two column groups of 16
    // share each tile); S0T sits after the strip partials in the scratch (16-byte aligned:
    // the partials are 1 KB x n)
    const int64_t npad = (int64_t)ntiles * kGT;
    double* S0T = part + (size_t)kGrmChunks * kGCMax * n;
    for (int c = 0; c < k; c += kGCMax) {
        const int kk = std::min(kGCMax, k - c);
        // 16 columns; 16 + 4 or 16 + 8 (4-column MFMA groups); two groups of 16
        const int cols = grm_padded_cols(kk);
        const unsigned gt = (unsigned)std::min<int64_t>((npad * cols + 255) / 256, 1024);
        k_grm_s0t<<<gt, 256, 0, s>>>(S0 + c * ldn, ldn, n, kk, cols, npad, S0T, skip);
        HIP_CHECK(hipGetLastError());
        if (cols == kGC)
            k_grm_tiles<1, 0><<<g, 256, 0, s>>>(tiles, n, S0T, kk, ntiles, part, skip);
        else if (cols == kGC + 4)
            k_grm_tiles<1, 1><<<g, 256, 0, s>>>(tiles, n, S0T, kk, ntiles, part, skip);
        else if (cols == kGC + 8)
            k_grm_tiles<1, 2><<<g, 256, 0, s>>>(tiles, n, S0T, kk, ntiles, part, skip);
        else
            k_grm_tiles<2, 0><<<g, 512, 0, s>>>(tiles, n, S0T, kk, ntiles, part, skip);
        HIP_CHECK(hipGetLastError());
        k_grm_finish<<<dim3(gf, kk), 256, 0, s>>>(part, n, kGCMax, theta_cols + c * 5, GZ + c * ldg, ldg,
                                                  skip);
        HIP_CHECK(hipGetLastError());
    }
}

size_t grm_sym_part_bytes(int32_t n, int k) {
    const size_t T = (size_t)(n + kGT - 1) / kGT;
    const size_t kc = (size_t)k;
    // row partials, column partials (k columns each), S0T (rows of 4 or 8 doubles)
    return sizeof(double) * (kc * ((size_t)kSymChunks * (size_t)n + T * (T + 1) / 2 * kGT) +
                             T * kGT * (k <= 4 ? 4 : 8));
}

void launch_grm_product_sym(hipStream_t s, const float* tiles, int32_t n, const double* S0, int64_t ldn, int k,
                            const double* theta_cols, double* GZ, int64_t ldg, const int* skip, double* scratch) {
    if (n == 0) return;
    const int32_t ntiles = (n + kGT - 1) / kGT;
    const int c4 = k <= 4 ? 4 : 8;
    const int64_t T = ntiles;
    double* rowp = scratch;
    double* colp = rowp + (size_t)kSymChunks * k * n;
    double* S0T = colp + (size_t)(T * (T + 1) / 2) * kGT * k;
    const dim3 g(ntiles, kSymChunks), gf(ntiles, k);
    if (k <= 2) {
        // one or two columns: the vector-ALU form (lane = tile column, the x values in
        // registers) streams as fast as the MFMA one and needs no transposed S0
        if (k == 1)
            k_grm_sym<1><<<g, 256, 0, s>>>(tiles, n, ntiles, S0, ldn, k, rowp, colp, skip);
        else
            k_grm_sym<2><<<g, 256, 0, s>>>(tiles, n, ntiles, S0, ldn, k, rowp, colp, skip);
    } else {
        const int64_t npad = T * kGT;
        const unsigned gt = (unsigned)std::min<int64_t>((npad * c4 + 255) / 256, 1024);
        k_grm_s0t<<<gt, 256, 0, s>>>(S0, ldn, n, k, c4, npad, S0T, skip);
        HIP_CHECK(hipGetLastError());
        if (k <= 4)
            k_grm_symx<4><<<g, 256, 0, s>>>(tiles, n, ntiles, S0T, k, rowp, colp, skip);
        else
            k_grm_symx<8><<<g, 256, 0, s>>>(tiles, n, ntiles, S0T, k, rowp, colp, skip);
    }
    HIP_CHECK(hipGetLastError());
    k_grm_sym_finish<<<gf, 256, 0, s>>>(rowp, colp, n, ntiles, k, theta_cols, GZ, ldg, skip);
    HIP_CHECK(hipGetLastError());
}

int grm_padded_cols(int k) {
    int cols = 0;
    for (int c = 0; c < k; c += kGCMax) {
        const int kk = std::min(kGCMax, k - c);
        cols += kk <= kGC ? kGC : kk <= kGC + 4 ? kGC + 4 : kk <= kGC + 8 ? kGC + 8 : 2 * kGC;
    }
    return cols;
}

size_t grm_product_part_bytes(int32_t n) {
    const size_t npad = (size_t)(n + kGT - 1) / kGT * kGT;
    return sizeof(double) * ((size_t)kGrmChunks * kGCMax * (size_t)n + npad * kGCMax);
}

void launch_grm_scatter(hipStream_t s, const double* Gc, int64_t ldg, int32_t n, int64_t j0, int k,
                        float* tiles) {
    if (k <= 0) return;
    const int64_t rows = n - j0;
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((rows + 255) / 256, 1024));
    k_grm_scatter<<<dim3(gx, k), 256, 0, s>>>(Gc, ldg, n, j0, k, tiles);
    HIP_CHECK(hipGetLastError());
}

}  // namespace globe
