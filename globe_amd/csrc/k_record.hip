// Record-space kernels: CSR segmented sums over an individual's records, the
// record-space output of V x, the per-individual Woodbury preconditioner, and the
// fused PCG vector updates with deterministic two-stage reductions (per-block
// partials in a fixed slot, then one ordered pass).
//
// Layout: record-space matrices column-major [k][nrec] (leading dimension ld);
// individual-space matrices [k][n]; theta_cols is [k][5] = per column
// {sig2e, sig2g, sig2b0, sigb01, sig2b1}.
#include <algorithm>

#include "globe_internal.h"

namespace globe {

namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum over a 256-thread block, result valid in thread 0. Fixed order: deterministic.
__device__ inline double block_sum(double v) {
    __shared__ double sh[kBlock / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) r += sh[i];
    return r;
}

// Start of column c of a record-space array with column stride ld: its own pool offset
// when the columns are pooled, else c ld.
__device__ inline int64_t col_start(const RecMeta& m, int c, int64_t ld) { return m.col_off ? m.col_off[c] : c * ld; }

// A tile's records (at most kTileRecords = kTileIters x kBlock) go to LDS: each thread
// issues all of its loads before its LDS stores, so the tile costs one memory round trip
// rather than one per record a thread copies.
constexpr int kTileIters = kTileRecords / kBlock;
static_assert(kTileIters * kBlock == kTileRecords, "tile records spread evenly over the block");
__device__ inline void stage_tile(const double* __restrict__ a, const double* __restrict__ b, int nr,
                                  double* __restrict__ sa, double* __restrict__ sb) {
    double va[kTileIters], vb[kTileIters];
#pragma unroll
    for (int i = 0; i < kTileIters; ++i) {
        const int q = threadIdx.x + kBlock * i;
        va[i] = q < nr ? a[q] : 0.0;
        vb[i] = q < nr ? b[q] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kTileIters; ++i) {
        const int q = threadIdx.x + kBlock * i;
        if (q < nr) {
            sa[q] = va[i];
            sb[q] = vb[i];
        }
    }
}

// ---------------------------------------------------------------------------
// S0[c][i] = sum_{r in block} x[c][r], S1[c][i] = sum t_r x[c][r]
// (PCGsolver.cpp:1119-1125 scatter, and the t-weighted sum the factorized time
// term needs instead of the n_b^2 loop at :1237-1254).
__global__ __launch_bounds__(kBlock) void k_seg_sums(int32_t nblk, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ ind,
                                                     const double* __restrict__ t,
                                                     const double* __restrict__ x, int64_t ldx,
                                                     double* __restrict__ S0,
                                                     double* __restrict__ S1, int64_t ldn,
                                                     const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double* xc = x + c * ldx;
    for (int64_t b = blockIdx.x * (int64_t)kBlock + threadIdx.x; b < nblk;
         b += (int64_t)gridDim.x * kBlock) {
        double s0 = 0.0, s1 = 0.0;
        for (int64_t r = off[b]; r < off[b + 1]; ++r) {
            double v = xc[r];
            s0 += v;
            s1 += t[r] * v;
        }
        const int32_t i = ind[b];
        S0[c * ldn + i] = s0;
        if (S1) S1[c * ldn + i] = s1;
    }
}

// ---------------------------------------------------------------- PCG control
// PCG per PCGsolver.cpp:1343-1443 / varCompEst.cpp:482-527 with every scalar on
// the device; once `done` is set every later kernel of the solve is a no-op, so the
// host may enqueue iterations ahead of the convergence test.
// sums[c] = sum_j part[c][j] for c < k: one wave per column; each lane loads its up to
// kMaxParts / 64 partials at once (so the loads are in flight together), adds them in
// the order j = lane, lane + 64, ... (slots j >= nparts read as 0), then a fixed
// shuffle tree (deterministic). Ends with a barrier.
constexpr int kPcgThreads = 1024;  // one-block PCG kernels: 16 waves, so k <= 64 columns take <= 4 rounds
// The sum of column c's partials by one wave (valid in lane 0): lane l adds slots
// l, l + 64, ... in order, then the shuffle tree.
__device__ inline double wave_col_sum(const double* __restrict__ part, int nparts, int c) {
    const int lane = threadIdx.x & 63;
    constexpr int kPer = kMaxParts / 64;
    double x[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int j = lane + 64 * i;
        x[i] = j < nparts ? part[c * kMaxParts + j] : 0.0;
    }
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) v += x[i];
    return wave_sum(v);
}
__device__ void wave_col_sums(const double* __restrict__ part, int nparts, int k, double* sums) {
    const int lane = threadIdx.x & 63;
    for (int c = threadIdx.x >> 6; c < k; c += blockDim.x >> 6) {
        const double v = wave_col_sum(part, nparts, c);
        if (lane == 0) sums[c] = v;
    }
    __syncthreads();
}

// alpha_c = r.z / p.Vp of the active columns (the p.Vp partials of k columns)
__device__ void pcg_alpha_body(PcgState* st, const double* __restrict__ part_pvp, int nparts, int k) {
    __shared__ double sums[kMaxCols];
    wave_col_sums(part_pvp, nparts, k, sums);
    const int c = threadIdx.x;
    if (c >= k) return;
    st->alpha[c] = st->active[c] ? st->rs[c] / sums[c] : 0.0;
    if (st->hist && st->active[c] && st->iters < st->hist_cap)
        st->hist[(1 + 2 * (int64_t)st->iters) * kMaxCols + c] = st->alpha[c];
}

// Convergence test after the x/r update of one iteration.
//   rule 0 (PCGsolver.cpp:1388-1415): max ratio over columns with |b| > 1e-12 below tol;
//          stall exit when the max ratio improves by < 5e-3 of its last value for 10
//          consecutive iterations.
//   rule 1 (varCompEst.cpp:505): every column |r| < tol |b|.
//   rule 2: per-column freeze; done when no column is active.
__device__ void pcg_check_body(PcgState* st, const double* __restrict__ part_rr, int nparts, int k) {
    __shared__ double nr_s[kMaxCols];
    wave_col_sums(part_rr, nparts, k, nr_s);
    if (threadIdx.x < k) nr_s[threadIdx.x] = sqrt(nr_s[threadIdx.x]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int it = st->iters + 1;
    int done = 0;
    double maxr = 0.0;
    if (st->rule == GLOBE_RULE_MAX_RATIO) {
        for (int j = 0; j < k; ++j) {
            const double nb = st->bnorm[j];
            const double ratio = nb > 1e-12 ? nr_s[j] / nb : 0.0;
            st->ratio[j] = ratio;
            if (ratio > maxr) maxr = ratio;
        }
        if (maxr < st->tol) {
            done = 1;
        } else if (st->stagnation) {
            const double impr = st->last - maxr, req = 5e-3 * st->last;
            st->stag = impr < req ? st->stag + 1 : 0;
            st->last = maxr;
            if (st->stag >= 10) {
                done = 1;
                st->stalled = 1;
            }
        } else {
            st->last = maxr;
        }
        for (int j = 0; j < k; ++j) st->col_iters[j] = it;
    } else if (st->rule == GLOBE_RULE_ALL_COLUMNS) {
        int all = 1;
        for (int j = 0; j < k; ++j) {
            const double nb = st->bnorm[j];
            if (!(nr_s[j] < st->tol * nb)) all = 0;
            const double ratio = nb > 0.0 ? nr_s[j] / nb : 0.0;
            st->ratio[j] = ratio;
            if (ratio > maxr) maxr = ratio;
            st->col_iters[j] = it;
        }
        done = all;
    } else {
        int any = 0;
        for (int j = 0; j < k; ++j) {
            const double nb = st->bnorm[j];
            const double ratio = nb > 1e-12 ? nr_s[j] / nb : 0.0;
            if (st->active[j]) {
                st->ratio[j] = nb > 0.0 ? nr_s[j] / nb : 0.0;
                st->col_iters[j] = it;
                if (nr_s[j] < st->tol * nb) st->active[j] = 0;
            }
            if (ratio > maxr) maxr = ratio;
            any |= st->active[j];
        }
        done = !any;
    }
    st->max_ratio = maxr;
    st->iters = it;
    if (it >= st->max_iter) done = 1;
    st->done = done;
}

// beta_c = r.z_new / r.z_old of the active columns (the r.z partials of k columns)
__device__ void pcg_beta_body(PcgState* st, const double* __restrict__ part_rz, int nparts, int k) {
    __shared__ double sums[kMaxCols];
    wave_col_sums(part_rz, nparts, k, sums);
    const int c = threadIdx.x;
    if (c >= k || !st->active[c]) return;
    const double s = sums[c];
    st->beta[c] = s / st->rs[c];
    st->rs[c] = s;
    // iters already counts this iteration
    if (st->hist && st->iters - 1 < st->hist_cap)
        st->hist[(2 + 2 * (int64_t)(st->iters - 1)) * kMaxCols + c] = st->beta[c];
}

// y[c][r] = GZ_i + sig2b0 S0_i + sigb01 (t S0_i + S1_i) + sig2b1 t S1_i + sig2e x_r
// (the factorized form of PCGsolver.cpp:1220-1265). Optional fused partial of
// sum_r x[c][r] y[c][r] (the p . Vp of PCG, PCGsolver.cpp:1381).
__global__ __launch_bounds__(kBlock) void k_record_out(
    RecMeta m, const double* __restrict__ x, int64_t ldx, const double* __restrict__ GZ,
    const double* __restrict__ S0, const double* __restrict__ S1, int64_t ldn,
    const double* __restrict__ th, double* __restrict__ y, int64_t ldy, double* part_dot,
    const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double se = th[c * 5 + 0], sb0 = th[c * 5 + 2], sb01 = th[c * 5 + 3],
                 sb1 = th[c * 5 + 4];
    const int32_t* indiv = m.indiv + col_start(m, c, m.rec_ld);
    const double* t = m.times + col_start(m, c, m.rec_ld);
    const int64_t nvalid = m.nrec_col ? m.nrec_col[c] : m.nrec;
    const double* xc = x + col_start(m, c, ldx);
    double* yc = y + col_start(m, c, ldy);
    // pooled columns hold exactly their own records; strided ones are padded to m.nrec
    const int64_t nloop = m.col_off ? nvalid : m.nrec;
    const double* g = GZ + c * ldn;
    const double* s0 = S0 + c * ldn;
    const double* s1 = S1 + c * ldn;
    double acc = 0.0;
    // the thread's records r0, r0 + stride, ... in groups of 4: the individual indices of
    // a group are loaded before its gathers (the same records in the same order)
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    constexpr int kU = 4;
    for (int64_t r0 = blockIdx.x * (int64_t)kBlock + threadIdx.x; r0 < nloop; r0 += kU * stride) {
        int32_t ii[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int64_t r = r0 + u * stride;
            ii[u] = r < nvalid ? indiv[r] : 0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int64_t r = r0 + u * stride;
            if (r >= nloop) break;
            double v = 0.0;  // records past a column's own count are padding
            if (r < nvalid) {
                const int32_t i = ii[u];
                const double tr = t[r], a = s0[i], b = s1[i], xv = xc[r];
                v = g[i] + sb0 * a + sb01 * (tr * a + b) + sb1 * tr * b + se * xv;
                acc += xv * v;
            }
            yc[r] = v;
        }
    }
    if (part_dot) {
        double s = block_sum(acc);
        if (threadIdx.x == 0) part_dot[c * kMaxParts + blockIdx.x] = s;
    }
}

// Tiled forms of the two per-block kernels: a workgroup stages the records of a tile
// of whole individual blocks in LDS with coalesced loads, one thread per block sums its
// records from LDS in record order (so the block sums equal the thread-per-block
// kernels' bit for bit), and the preconditioner writes z back one thread per record.
// With pst (inside PCG), x is the search direction p and is first updated in place,
// p = z + beta p for the active columns (PCGsolver.cpp:1427), so the sums see the new p.
__global__ __launch_bounds__(kBlock) void k_seg_sums_tiled(RecMeta m, double* __restrict__ x,
                                                           int64_t ldx, double* __restrict__ S0,
                                                           double* __restrict__ S1, int64_t ldn,
                                                           const int* skip, const double* __restrict__ z,
                                                           const PcgState* pst) {
    if (skip && *skip) return;
    extern __shared__ double sm[];
    double* sx = sm;
    double* st = sm + m.cap;
    const int c = blockIdx.y;
    const bool upd = pst && pst->active[c];
    const double beta = upd ? pst->beta[c] : 0.0;
    double* xc = x + col_start(m, c, ldx);
    const double* zc = upd ? z + col_start(m, c, ldx) : nullptr;
    const double* t = m.times + col_start(m, c, m.rec_ld);
    const int64_t* off = m.blk_off + col_start(m, c, m.blk_ld);
    const int32_t* ind = m.blk_ind + col_start(m, c, m.blk_ld);
    const int32_t* tile_blk = m.tile_blk + col_start(m, c, m.tile_ld);
    const int32_t ntiles = m.ntiles_col ? m.ntiles_col[c] : m.ntiles;
    for (int32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int32_t b0 = tile_blk[tile], b1 = tile_blk[tile + 1];
        const int64_t r0 = off[b0];
        const int nr = (int)(off[b1] - r0);
        __syncthreads();
        if (upd) {
            double vp[kTileIters], vz[kTileIters], vt[kTileIters];
#pragma unroll
            for (int i = 0; i < kTileIters; ++i) {
                const int q = threadIdx.x + kBlock * i;
                vp[i] = q < nr ? xc[r0 + q] : 0.0;
                vz[i] = q < nr ? zc[r0 + q] : 0.0;
                vt[i] = q < nr ? t[r0 + q] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < kTileIters; ++i) {
                const int q = threadIdx.x + kBlock * i;
                if (q < nr) {
                    const double pv = vz[i] + beta * vp[i];
                    xc[r0 + q] = pv;
                    sx[q] = pv;
                    st[q] = vt[i];
                }
            }
        } else {
            stage_tile(xc + r0, t + r0, nr, sx, st);
        }
        __syncthreads();
        for (int32_t b = b0 + threadIdx.x; b < b1; b += kBlock) {
            const int q0 = (int)(off[b] - r0), q1 = (int)(off[b + 1] - r0);
            double s0 = 0.0, s1 = 0.0;
            for (int q = q0; q < q1; ++q) {
                const double v = sx[q];
                s0 += v;
                s1 += st[q] * v;
            }
            const int32_t i = ind[b];
            S0[c * ldn + i] = s0;
            if (S1) S1[c * ldn + i] = s1;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_precond_tiled(
    RecMeta m, const double* __restrict__ gdiag, const double* __restrict__ th,
    const double* __restrict__ r, int64_t ldr, double* __restrict__ z, int64_t ldz,
    double* part_rz, double* part_rr, const PcgState* stt, int tilde_v) {
    if (stt && stt->done) return;
    const int c = blockIdx.y;
    if (stt && !stt->active[c]) {
        // frozen column: keep the partial slots defined
        if (threadIdx.x == 0) {
            if (part_rz) part_rz[c * kMaxParts + blockIdx.x] = 0.0;
            if (part_rr) part_rr[c * kMaxParts + blockIdx.x] = 0.0;
        }
        return;
    }
    extern __shared__ double sm[];
    double* sr = sm;               // r of the tile's records
    double* st = sm + m.cap;       // t of the tile's records
    double* sc0 = sm + 2 * m.cap;  // per block (at most kTileBlocks): Woodbury coefficients
    double* sc1 = sc0 + kTileBlocks;
    const double se = th[c * 5 + 0], sg = th[c * 5 + 1], sb0 = th[c * 5 + 2],
                 sb01 = th[c * 5 + 3], sb1 = th[c * 5 + 4];
    const double* rc = r + col_start(m, c, ldr);
    double* zc = z + col_start(m, c, ldz);
    const double* t = m.times + col_start(m, c, m.rec_ld);
    const int32_t* rec_blk = m.rec_blk + col_start(m, c, m.rec_ld);
    const int64_t* off = m.blk_off + col_start(m, c, m.blk_ld);
    const int32_t* ind = m.blk_ind + col_start(m, c, m.blk_ld);
    const int32_t* tile_blk = m.tile_blk + col_start(m, c, m.tile_ld);
    const int32_t ntiles = m.ntiles_col ? m.ntiles_col[c] : m.ntiles;
    const double inv_se = 1.0 / se;
    double arz = 0.0, arr = 0.0;
    for (int32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int32_t b0 = tile_blk[tile], b1 = tile_blk[tile + 1];
        const int64_t r0 = off[b0];
        const int nr = (int)(off[b1] - r0);
        // the blocks of this thread's records, loaded with the tile (used after the barrier)
        int32_t rb[kTileIters];
#pragma unroll
        for (int i = 0; i < kTileIters; ++i) {
            const int q = threadIdx.x + kBlock * i;
            rb[i] = q < nr ? rec_blk[r0 + q] : 0;
        }
        __syncthreads();
        stage_tile(rc + r0, t + r0, nr, sr, st);
        __syncthreads();
        for (int32_t b = b0 + threadIdx.x; b < b1; b += kBlock) {
            const int q0 = (int)(off[b] - r0), q1 = (int)(off[b + 1] - r0);
            double a01 = 0.0, a11 = 0.0, w0 = 0.0, w1 = 0.0;
            for (int q = q0; q < q1; ++q) {
                const double tq = st[q], rv = sr[q];
                a01 += tq;
                a11 += tq * tq;
                w0 += rv;
                w1 += tq * rv;
            }
            const double a00 = (double)(q1 - q0);
            const double gv = sg * (tilde_v ? 1.0 : gdiag[ind[b]]);
            const double c00 = gv + sb0, c01 = sb01, c11 = sb1;
            const double m00 = se + c00 * a00 + c01 * a01;
            const double m01 = c00 * a01 + c01 * a11;
            const double m10 = c01 * a00 + c11 * a01;
            const double m11 = se + c01 * a01 + c11 * a11;
            const double v0 = c00 * w0 + c01 * w1;
            const double v1 = c01 * w0 + c11 * w1;
            const double det = m00 * m11 - m01 * m10;
            sc0[b - b0] = (m11 * v0 - m01 * v1) / det;
            sc1[b - b0] = (m00 * v1 - m10 * v0) / det;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kTileIters; ++i) {
            const int q = threadIdx.x + kBlock * i;
            if (q >= nr) break;
            const int bl = rb[i] - b0;
            const double rv = sr[q];
            const double zv = (rv - sc0[bl] - sc1[bl] * st[q]) * inv_se;
            zc[r0 + q] = zv;
            arz += rv * zv;
            arr += rv * rv;
        }
    }
    if (part_rz) {
        double s = block_sum(arz);
        if (threadIdx.x == 0) part_rz[c * kMaxParts + blockIdx.x] = s;
    }
    if (part_rr) {
        double s = block_sum(arr);
        if (threadIdx.x == 0) part_rr[c * kMaxParts + blockIdx.x] = s;
    }
}

// One PCG step after V p, fused: alpha = r.z / p.Vp (every workgroup reduces the p.Vp
// partials of its column itself, in the order of wave_col_sums, so all of them hold the
// same alpha; workgroup 0 of the column records it), x += alpha p and r -= alpha Vp per
// active column (PCGsolver.cpp:1381-1386), then z = Q^-1 r by the preconditioner of
// k_precond_tiled on the updated r, with the r.r and r.z partials of the new r. Frozen
// columns (per-column rule) only give their r.r partials. The convergence test and beta
// follow in one single-block kernel (k_pcg_check_beta), the only grid-wide reduction of
// an iteration: a column that the test freezes has its z computed here but never used.
__global__ __launch_bounds__(kBlock) void k_pcg_xr_precond(
    RecMeta m, const double* __restrict__ gdiag, const double* __restrict__ th, PcgState* stt,
    const double* __restrict__ part_pvp, int np, double* __restrict__ x, int64_t ldx,
    double* __restrict__ r, const double* __restrict__ p, const double* __restrict__ vp, int64_t ld,
    double* __restrict__ z, double* part_rz, double* part_rr) {
    if (stt->done) return;
    const int c = blockIdx.y;
    const bool act = stt->active[c] != 0;
    __shared__ double alpha_s;
    if (threadIdx.x < 64) {
        const double sum = wave_col_sum(part_pvp, np, c);
        if (threadIdx.x == 0) {
            const double a = act ? stt->rs[c] / sum : 0.0;
            alpha_s = a;
            if (blockIdx.x == 0) {
                stt->alpha[c] = a;
                if (stt->hist && act && stt->iters < stt->hist_cap)
                    stt->hist[(1 + 2 * (int64_t)stt->iters) * kMaxCols + c] = a;
            }
        }
    }
    __syncthreads();
    const double alpha = alpha_s;
    extern __shared__ double sm[];
    double* sr = sm;               // r of the tile's records (updated)
    double* st = sm + m.cap;       // t of the tile's records
    double* sc0 = sm + 2 * m.cap;  // per block (at most kTileBlocks): Woodbury coefficients
    double* sc1 = sc0 + kTileBlocks;
    const double se = th[c * 5 + 0], sg = th[c * 5 + 1], sb0 = th[c * 5 + 2],
                 sb01 = th[c * 5 + 3], sb1 = th[c * 5 + 4];
    double* rc = r + col_start(m, c, ld);
    double* xc = x + col_start(m, c, ldx);
    const double* pc = p + col_start(m, c, ld);
    const double* vc = vp + col_start(m, c, ld);
    double* zc = z + col_start(m, c, ld);
    const double* t = m.times + col_start(m, c, m.rec_ld);
    const int32_t* rec_blk = m.rec_blk + col_start(m, c, m.rec_ld);
    const int64_t* off = m.blk_off + col_start(m, c, m.blk_ld);
    const int32_t* ind = m.blk_ind + col_start(m, c, m.blk_ld);
    const int32_t* tile_blk = m.tile_blk + col_start(m, c, m.tile_ld);
    const int32_t ntiles = m.ntiles_col ? m.ntiles_col[c] : m.ntiles;
    const double inv_se = 1.0 / se;
    double arz = 0.0, arr = 0.0;
    for (int32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int32_t b0 = tile_blk[tile], b1 = tile_blk[tile + 1];
        const int64_t r0 = off[b0];
        const int nr = (int)(off[b1] - r0);
        if (!act) {
            for (int q = threadIdx.x; q < nr; q += kBlock) {
                const double rv = rc[r0 + q];
                arr += rv * rv;
            }
            continue;
        }
        // all of this thread's loads of the tile before any store
        int32_t rb[kTileIters];
        double vr[kTileIters], vv[kTileIters], vpp[kTileIters], vx[kTileIters], vt[kTileIters];
#pragma unroll
        for (int i = 0; i < kTileIters; ++i) {
            const int q = threadIdx.x + kBlock * i;
            const bool in = q < nr;
            rb[i] = in ? rec_blk[r0 + q] : 0;
            vr[i] = in ? rc[r0 + q] : 0.0;
            vv[i] = in ? vc[r0 + q] : 0.0;
            vpp[i] = in ? pc[r0 + q] : 0.0;
            vx[i] = in ? xc[r0 + q] : 0.0;
            vt[i] = in ? t[r0 + q] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kTileIters; ++i) {
            const int q = threadIdx.x + kBlock * i;
            if (q < nr) {
                const double xv = vx[i] + alpha * vpp[i];
                const double rv = vr[i] - alpha * vv[i];
                xc[r0 + q] = xv;
                rc[r0 + q] = rv;
                vr[i] = rv;
                sr[q] = rv;
                st[q] = vt[i];
                arr += rv * rv;
            }
        }
        __syncthreads();
        for (int32_t b = b0 + threadIdx.x; b < b1; b += kBlock) {
            const int q0 = (int)(off[b] - r0), q1 = (int)(off[b + 1] - r0);
            double a01 = 0.0, a11 = 0.0, w0 = 0.0, w1 = 0.0;
            for (int q = q0; q < q1; ++q) {
                const double tq = st[q], rv = sr[q];
                a01 += tq;
                a11 += tq * tq;
                w0 += rv;
                w1 += tq * rv;
            }
            const double a00 = (double)(q1 - q0);
            const double gv = sg * gdiag[ind[b]];
            const double c00 = gv + sb0, c01 = sb01, c11 = sb1;
            const double m00 = se + c00 * a00 + c01 * a01;
            const double m01 = c00 * a01 + c01 * a11;
            const double m10 = c01 * a00 + c11 * a01;
            const double m11 = se + c01 * a01 + c11 * a11;
            const double v0 = c00 * w0 + c01 * w1;
            const double v1 = c01 * w0 + c11 * w1;
            const double det = m00 * m11 - m01 * m10;
            sc0[b - b0] = (m11 * v0 - m01 * v1) / det;
            sc1[b - b0] = (m00 * v1 - m10 * v0) / det;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kTileIters; ++i) {
            const int q = threadIdx.x + kBlock * i;
            if (q >= nr) break;
            const int bl = rb[i] - b0;
            const double rv = vr[i];
            const double zv = (rv - sc0[bl] - sc1[bl] * vt[i]) * inv_se;
            zc[r0 + q] = zv;
            arz += rv * zv;
        }
    }
    double s = block_sum(arz);
    if (threadIdx.x == 0) part_rz[c * kMaxParts + blockIdx.x] = s;
    s = block_sum(arr);
    if (threadIdx.x == 0) part_rr[c * kMaxParts + blockIdx.x] = s;
}

// z = Q_i^-1 r per individual block by the 2x2 Woodbury form of
// Q_i = sig2e I + U_i C_i U_i^T, U_i = [1, t], C_i = [[sig2g G_ii + sig2b0, sigb01],
// [sigb01, sig2b1]] -- the same block the reference inverts explicitly
// (PCGsolver.cpp:1026-1045, applied at :1060-1093):
//   w = U^T r, c = (sig2e I + C U^T U)^-1 C w, z = (r - U c) / sig2e.
// tilde_v = 1 uses G_ii := 1 (tilde-V, varCompEst.cpp:839-858).
// Fused partials: sum r.z and sum r.r per column.
__global__ __launch_bounds__(kBlock) void k_precond(
    int32_t nblk, const int64_t* __restrict__ off, const int32_t* __restrict__ ind,
    const double* __restrict__ t, const double* __restrict__ gdiag, const double* __restrict__ th,
    const double* __restrict__ r, int64_t ldr, double* __restrict__ z, int64_t ldz,
    double* part_rz, double* part_rr, const PcgState* st, int tilde_v) {
    if (st && st->done) return;
    const int c = blockIdx.y;
    if (st && !st->active[c]) {
        // frozen column: keep the partial slots defined
        if (threadIdx.x == 0) {
            if (part_rz) part_rz[c * kMaxParts + blockIdx.x] = 0.0;
            if (part_rr) part_rr[c * kMaxParts + blockIdx.x] = 0.0;
        }
        return;
    }
    const double se = th[c * 5 + 0], sg = th[c * 5 + 1], sb0 = th[c * 5 + 2],
                 sb01 = th[c * 5 + 3], sb1 = th[c * 5 + 4];
    const double* rc = r + c * ldr;
    double* zc = z + c * ldz;
    const double inv_se = 1.0 / se;
    double arz = 0.0, arr = 0.0;
    for (int64_t b = blockIdx.x * (int64_t)kBlock + threadIdx.x; b < nblk;
         b += (int64_t)gridDim.x * kBlock) {
        const int64_t r0 = off[b], r1 = off[b + 1];
        double a01 = 0.0, a11 = 0.0, w0 = 0.0, w1 = 0.0;
        for (int64_t q = r0; q < r1; ++q) {
            const double tq = t[q], rv = rc[q];
            a01 += tq;
            a11 += tq * tq;
            w0 += rv;
            w1 += tq * rv;
        }
        const double a00 = (double)(r1 - r0);
        const double gv = sg * (tilde_v ? 1.0 : gdiag[ind[b]]);
        const double c00 = gv + sb0, c01 = sb01, c11 = sb1;
        // M2 = se I + C A (A = U^T U symmetric)
        const double m00 = se + c00 * a00 + c01 * a01;
        const double m01 = c00 * a01 + c01 * a11;
        const double m10 = c01 * a00 + c11 * a01;
        const double m11 = se + c01 * a01 + c11 * a11;
        const double v0 = c00 * w0 + c01 * w1;
        const double v1 = c01 * w0 + c11 * w1;
        const double det = m00 * m11 - m01 * m10;
        const double cc0 = (m11 * v0 - m01 * v1) / det;
        const double cc1 = (m00 * v1 - m10 * v0) / det;
        for (int64_t q = r0; q < r1; ++q) {
            const double rv = rc[q];
            const double zv = (rv - cc0 - cc1 * t[q]) * inv_se;
            zc[q] = zv;
            arz += rv * zv;
            arr += rv * rv;
        }
    }
    if (part_rz) {
        double s = block_sum(arz);
        if (threadIdx.x == 0) part_rz[c * kMaxParts + blockIdx.x] = s;
    }
    if (part_rr) {
        double s = block_sum(arr);
        if (threadIdx.x == 0) part_rr[c * kMaxParts + blockIdx.x] = s;
    }
}

// log|Q_i| of the preconditioner blocks by Sylvester's determinant identity:
// |sig2e I_n + U C U^T| = sig2e^(n - 2) |sig2e I_2 + C U^T U| (the 2x2 matrix k_precond
// inverts); per-workgroup partial sums into part[blockIdx.x].
__global__ __launch_bounds__(kBlock) void k_logdet_q(int32_t nblk, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ ind,
                                                     const double* __restrict__ t,
                                                     const double* __restrict__ gdiag,
                                                     const double* __restrict__ th, double* part) {
    const double se = th[0], sg = th[1], sb0 = th[2], sb01 = th[3], sb1 = th[4];
    const double lse = log(se);
    double acc = 0.0;
    for (int64_t b = blockIdx.x * (int64_t)kBlock + threadIdx.x; b < nblk;
         b += (int64_t)gridDim.x * kBlock) {
        const int64_t r0 = off[b], r1 = off[b + 1];
        double a01 = 0.0, a11 = 0.0;
        for (int64_t q = r0; q < r1; ++q) {
            a01 += t[q];
            a11 += t[q] * t[q];
        }
        const double a00 = (double)(r1 - r0);
        const double c00 = sg * gdiag[ind[b]] + sb0, c01 = sb01, c11 = sb1;
        const double m00 = se + c00 * a00 + c01 * a01;
        const double m01 = c00 * a01 + c01 * a11;
        const double m10 = c01 * a00 + c11 * a01;
        const double m11 = se + c01 * a01 + c11 * a11;
        acc += (a00 - 2.0) * lse + log(m00 * m11 - m01 * m10);
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_dot_parts(int64_t nrec, const double* __restrict__ a,
                                                      int64_t lda, const double* __restrict__ b,
                                                      int64_t ldb, double* part) {
    const int c = blockIdx.y;
    const double* ac = a + c * lda;
    const double* bc = b + c * ldb;
    double acc = 0.0;
    for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < nrec;
         r += (int64_t)gridDim.x * kBlock)
        acc += ac[r] * bc[r];
    double s = block_sum(acc);
    if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = s;
}

// out[c] = sum_j part[c][j], j ascending (one thread per column).
// out[c] = sum of the nparts partials of column c: one block per column, thread t adds
// partials t, t + 256, ... in order, then a fixed-order block sum (deterministic; a
// thread per column would chain up to 1024 dependent loads).
__global__ __launch_bounds__(kBlock) void k_reduce_parts(const double* __restrict__ part, int nparts,
                                                         double* out, const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.x;
    double s = 0.0;
    for (int j = threadIdx.x; j < nparts; j += kBlock) s += part[c * kMaxParts + j];
    s = block_sum(s);
    if (threadIdx.x == 0) out[c] = s;
}

// A^T B partials: part[(a*kb + b)][block] = column a of A . column (a bsa + b) of B
// (bsa = 0: the full gram; bsa = kb: column a of A against its own group of kb columns)
__global__ __launch_bounds__(kBlock) void k_gram_parts(int64_t nrec, const double* __restrict__ A,
                                                       int64_t lda, const double* __restrict__ B,
                                                       int64_t ldb, int kb, int bsa, double* part) {
    const int a = blockIdx.y / kb, bb = blockIdx.y % kb;
    const double* ac = A + a * lda;
    const double* bc = B + (int64_t)(a * bsa + bb) * ldb;
    double acc = 0.0;
    for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < nrec;
         r += (int64_t)gridDim.x * kBlock)
        acc += ac[r] * bc[r];
    double s = block_sum(acc);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.y * kMaxParts + blockIdx.x] = s;
}

// y[:, j] -= sum_q x[:, q] coef[q * ky + j]   (V^-1 r = V^-1 y - V^-1 X beta, P f)
__global__ __launch_bounds__(kBlock) void k_axpy_cols(int64_t nrec, double* __restrict__ y,
                                                      int64_t ldy, const double* __restrict__ x,
                                                      int64_t ldx, int kx,
                                                      const double* __restrict__ coef, int ky) {
    const int j = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x; r < nrec;
         r += (int64_t)gridDim.x * kBlock) {
        double v = y[j * ldy + r];
        for (int q = 0; q < kx; ++q) v -= x[q * ldx + r] * coef[q * ky + j];
        y[j * ldy + r] = v;
    }
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg_init(PcgState* st, const double* __restrict__ part_bb,
                                                     const double* __restrict__ part_rr,
                                                     const double* __restrict__ part_rz, int nparts,
                                                     int k) {
    const int c = threadIdx.x;
    __shared__ double ratio_s[kMaxCols];
    __shared__ int act_s[kMaxCols];
    __shared__ double sbb[kMaxCols], srr[kMaxCols], srz[kMaxCols];
    wave_col_sums(part_bb, nparts, k, sbb);
    wave_col_sums(part_rr, nparts, k, srr);
    wave_col_sums(part_rz, nparts, k, srz);
    if (c < k) {
        const double bb = sbb[c], rr = srr[c], rz = srz[c];
        const double nb = sqrt(bb), nr = sqrt(rr);
        st->bnorm[c] = nb;
        st->rs[c] = rz;
        st->alpha[c] = 0.0;
        st->beta[c] = 0.0;
        st->col_iters[c] = 0;
        if (st->hist) st->hist[c] = rz;
        st->ratio[c] = nb > 0.0 ? nr / nb : 0.0;
        ratio_s[c] = nb > 1e-12 ? nr / nb : 0.0;
        int act = 1;
        if (st->rule == GLOBE_RULE_PER_COLUMN && (nb <= 1e-12 || nr < st->tol * nb)) act = 0;
        st->active[c] = act;
        act_s[c] = act;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double last = 0.0;
        int any = 0;
        for (int j = 0; j < k; ++j) {
            if (ratio_s[j] > last) last = ratio_s[j];
            any |= act_s[j];
        }
        st->last = last;
        st->max_ratio = last;
        st->iters = 0;
        st->stag = 0;
        st->stalled = 0;
        st->done = (st->rule == GLOBE_RULE_PER_COLUMN && !any) || st->max_iter <= 0;
    }
}

// Start (or restart) column j of a running block PCG from the partials of its
// first residual (column 0 of each partial array): |b|, |r|, r.z; the column becomes
// active unless it is already converged, and a block whose columns had all finished
// runs again. Used by the serving loop when a request joins between iterations.
__global__ __launch_bounds__(64) void k_pcg_admit(PcgState* st, int j, const double* __restrict__ part_bb,
                                                  int nparts_bb, const double* __restrict__ part_rr,
                                                  const double* __restrict__ part_rz, int nparts) {
    const int lane = threadIdx.x;
    double bb = 0.0, rr = 0.0, rz = 0.0;
    for (int q = lane; q < nparts_bb; q += 64) bb += part_bb[q];
    for (int q = lane; q < nparts; q += 64) {
        rr += part_rr[q];
        rz += part_rz[q];
    }
    bb = wave_sum(bb);
    rr = wave_sum(rr);
    rz = wave_sum(rz);
    if (lane != 0) return;
    const double nb = sqrt(bb), nr = sqrt(rr);
    st->bnorm[j] = nb;
    st->rs[j] = rz;
    st->alpha[j] = 0.0;
    st->beta[j] = 0.0;
    st->col_iters[j] = 0;
    st->ratio[j] = nb > 0.0 ? nr / nb : 0.0;
    const int act = (nb > 1e-12 && !(nr < st->tol * nb)) ? 1 : 0;
    st->active[j] = act;
    if (act) st->done = 0;
}

// Pooled columns (col_off set): column c starts at col_off[c] of every vector and holds
// col_cnt[c] records; otherwise it starts at c ld (c ldx for x) and holds nrec.
// Place column j of a pooled block (the serving loop): its pool offset, tile and record
// counts and variance components, passed by value so that no host buffer is read.
__global__ void k_set_column(int j, int64_t off, int32_t ntiles, int64_t nrec, double t0, double t1, double t2,
                             double t3, double t4, int64_t* col_off, int32_t* ntiles_col, int64_t* nrec_col,
                             double* theta) {
    if (threadIdx.x != 0) return;
    col_off[j] = off;
    ntiles_col[j] = ntiles;
    nrec_col[j] = nrec;
    theta[5 * j + 0] = t0;
    theta[5 * j + 1] = t1;
    theta[5 * j + 2] = t2;
    theta[5 * j + 3] = t3;
    theta[5 * j + 4] = t4;
}

// Column j stops iterating (the serving loop's iteration cap).
__global__ void k_pcg_deactivate(PcgState* st, int j) {
    if (threadIdx.x == 0) st->active[j] = 0;
}

// One-block reductions of PCG's dot partials: alpha after V p, the convergence test
// after the x/r update, beta after the preconditioner.
__global__ __launch_bounds__(kPcgThreads) void k_pcg_alpha(PcgState* st, const double* __restrict__ part_pvp,
                                                      int nparts, int k) {
    if (st->done) return;
    pcg_alpha_body(st, part_pvp, nparts, k);
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg_check(PcgState* st, const double* __restrict__ part_rr,
                                                      int nparts, int k) {
    if (st->done) return;
    pcg_check_body(st, part_rr, nparts, k);
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg_beta(PcgState* st, const double* __restrict__ part_rz,
                                                     int nparts, int k) {
    if (st->done) return;
    pcg_beta_body(st, part_rz, nparts, k);
}

// The convergence test, then beta if the solve goes on (one launch per iteration)
__global__ __launch_bounds__(kPcgThreads) void k_pcg_check_beta(PcgState* st, const double* __restrict__ part_rr,
                                                           int nparts_rr, const double* __restrict__ part_rz,
                                                           int nparts_rz, int k) {
    if (st->done) return;
    pcg_check_body(st, part_rr, nparts_rr, k);
    __syncthreads();
    if (st->done) return;
    pcg_beta_body(st, part_rz, nparts_rz, k);
}

__global__ __launch_bounds__(kBlock) void k_pcg_xr(PcgState* st, int64_t nrec, double* __restrict__ x,
                                                   int64_t ldx, double* __restrict__ r,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ vp, int64_t ld,
                                                   double* part_rr, const int64_t* __restrict__ col_off,
                                                   const int64_t* __restrict__ col_cnt) {
    if (st->done) return;
    const int c = blockIdx.y;
    const double a = st->alpha[c];
    const bool act = st->active[c] != 0;
    const int64_t o = col_off ? col_off[c] : c * ld, ox = col_off ? col_off[c] : c * ldx;
    const int64_t nc = col_cnt ? col_cnt[c] : nrec;
    double acc = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < nc;
         q += (int64_t)gridDim.x * kBlock) {
        double rv = r[o + q];
        if (act) {
            x[ox + q] += a * p[o + q];
            rv -= a * vp[o + q];
            r[o + q] = rv;
        }
        acc += rv * rv;
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) part_rr[c * kMaxParts + blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_pcg_p(PcgState* st, int64_t nrec, double* __restrict__ p,
                                                  const double* __restrict__ z, int64_t ld,
                                                  const int64_t* __restrict__ col_off,
                                                  const int64_t* __restrict__ col_cnt) {
    if (st->done) return;
    const int c = blockIdx.y;
    if (!st->active[c]) return;
    const double b = st->beta[c];
    const int64_t o = col_off ? col_off[c] : c * ld;
    const int64_t nc = col_cnt ? col_cnt[c] : nrec;
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < nc;
         q += (int64_t)gridDim.x * kBlock)
        p[o + q] = z[o + q] + b * p[o + q];
}

// bhat[2i] = sum v_r (sig2b0 + sigb01 t_r), bhat[2i+1] = sum v_r (sigb01 + sig2b1 t_r)
// (PCGsolver.cpp:1491-1504), one thread per individual: fixed summation order.
__global__ __launch_bounds__(kBlock) void k_bhat(int32_t nblk, const int64_t* __restrict__ off,
                                                 const int32_t* __restrict__ ind,
                                                 const double* __restrict__ t,
                                                 const double* __restrict__ v,
                                                 const double* __restrict__ th,
                                                 double* __restrict__ bhat) {
    const double sb0 = th[2], sb01 = th[3], sb1 = th[4];
    for (int64_t b = blockIdx.x * (int64_t)kBlock + threadIdx.x; b < nblk;
         b += (int64_t)gridDim.x * kBlock) {
        double u0 = 0.0, u1 = 0.0;
        for (int64_t q = off[b]; q < off[b + 1]; ++q) {
            u0 += v[q] * (sb0 + sb01 * t[q]);
            u1 += v[q] * (sb01 + sb1 * t[q]);
        }
        const int32_t i = ind[b];
        bhat[2 * i] = u0;
        bhat[2 * i + 1] = u1;
    }
}

// f_theta for ncols P-vectors (varCompEst.cpp:693-760): f[(5*j + q)][r], q = 0..4
// from py[j] and uhat[j].
__global__ __launch_bounds__(kBlock) void k_f_theta(int32_t nblk, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ ind,
                                                    const double* __restrict__ t,
                                                    const double* __restrict__ py, int64_t ldpy,
                                                    const double* __restrict__ uhat, int64_t ldu,
                                                    double inv_sg, double* __restrict__ f,
                                                    int64_t ldf) {
    const int j = blockIdx.y;
    const double* pj = py + j * ldpy;
    const double* uj = uhat + j * ldu;
    double* fj = f + (int64_t)5 * j * ldf;
    for (int64_t b = blockIdx.x * (int64_t)kBlock + threadIdx.x; b < nblk;
         b += (int64_t)gridDim.x * kBlock) {
        double w = 0.0, tw = 0.0;
        for (int64_t q = off[b]; q < off[b + 1]; ++q) {
            w += pj[q];
            tw += pj[q] * t[q];
        }
        const double u = uj[ind[b]] * inv_sg;
        for (int64_t q = off[b]; q < off[b + 1]; ++q) {
            const double tq = t[q];
            fj[0 * ldf + q] = pj[q];
            fj[1 * ldf + q] = u;
            fj[2 * ldf + q] = w;
            fj[3 * ldf + q] = tq * w + tw;
            fj[4 * ldf + q] = tw * tq;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_copy_cols(int64_t rows, const double* __restrict__ src,
                                                      int64_t lds, double* __restrict__ dst,
                                                      int64_t ldd) {
    const int c = blockIdx.y;
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < rows;
         q += (int64_t)gridDim.x * kBlock)
        dst[c * ldd + q] = src[c * lds + q];
}

// d[c][q] = a[c][q] - b[c][q]  (r = b - V x0)
__global__ __launch_bounds__(kBlock) void k_sub_cols(int64_t rows, const double* __restrict__ a,
                                                     int64_t lda, const double* __restrict__ b,
                                                     int64_t ldb, double* __restrict__ d,
                                                     int64_t ldd) {
    const int c = blockIdx.y;
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < rows;
         q += (int64_t)gridDim.x * kBlock)
        d[c * ldd + q] = a[c * lda + q] - b[c * ldb + q];
}

__global__ __launch_bounds__(kBlock) void k_fill_cols(int64_t rows, double* __restrict__ d,
                                                      int64_t ldd, double v) {
    const int c = blockIdx.y;
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < rows;
         q += (int64_t)gridDim.x * kBlock)
        d[c * ldd + q] = v;
}

__global__ __launch_bounds__(kBlock) void k_copy16(uint4* __restrict__ dst,
                                                   const uint4* __restrict__ src, int64_t n16) {
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < n16;
         q += (int64_t)gridDim.x * kBlock)
        dst[q] = src[q];
}

__global__ __launch_bounds__(kBlock) void k_fill(double* dst, int64_t count, double v) {
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < count;
         q += (int64_t)gridDim.x * kBlock)
        dst[q] = v;
}

// Zero the partial slots [from, to) of every column of up to three partial arrays
// (kMaxParts slots per column; null arrays are skipped): one launch for all of them.
__global__ __launch_bounds__(kBlock) void k_zero_slots(double* a, double* b, double* c, int k, int from, int to) {
    double* arr = blockIdx.y == 0 ? a : blockIdx.y == 1 ? b : c;
    if (!arr) return;
    for (int q = threadIdx.x; q < k * (to - from); q += kBlock) {
        const int j = q / (to - from), slot = from + q % (to - from);
        arr[j * kMaxParts + slot] = 0.0;
    }
}

}  // namespace

#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

int precond_parts(const RecMeta& m) {
    return m.cap > 0 ? std::max(1, std::min<int>(m.ntiles, kMaxParts)) : parts_for(m.nblk);
}

void launch_seg_sums(hipStream_t s, const RecMeta& m, const double* x, int64_t ldx, int k,
                     double* S0, double* S1, int64_t ldn, const int* skip) {
    if (m.cap > 0) {
        dim3 g(precond_parts(m), k);
        k_seg_sums_tiled<<<g, kBlock, sizeof(double) * 2 * m.cap, s>>>(m, const_cast<double*>(x), ldx, S0, S1,
                                                                        ldn, skip, nullptr, nullptr);
    } else {
        dim3 g(parts_for(m.nblk), k);
        k_seg_sums<<<g, kBlock, 0, s>>>(m.nblk, m.blk_off, m.blk_ind, m.times, x, ldx, S0, S1, ldn,
                                         skip);
    }
    LAUNCH_CHECK();
}

bool launch_seg_sums_pcg(hipStream_t s, const RecMeta& m, double* p, const double* z, int64_t ld, int k,
                         PcgState* st, double* S0, double* S1, int64_t ldn) {
    if (m.cap == 0) return false;
    dim3 g(precond_parts(m), k);
    k_seg_sums_tiled<<<g, kBlock, sizeof(double) * 2 * m.cap, s>>>(m, p, ld, S0, S1, ldn, &st->done, z, st);
    LAUNCH_CHECK();
    return true;
}

bool launch_pcg_xr_precond(hipStream_t s, const RecMeta& m, const double* gdiag, const double* theta_cols,
                           PcgState* st, const double* part_pvp, int np, double* x, int64_t ldx, double* r,
                           const double* p, const double* vp, int64_t ld, int k, double* z, double* part_rz,
                           double* part_rr) {
    if (m.cap == 0) return false;
    dim3 g(precond_parts(m), k);
    k_pcg_xr_precond<<<g, kBlock, sizeof(double) * (2 * m.cap + 2 * kTileBlocks), s>>>(
        m, gdiag, theta_cols, st, part_pvp, np, x, ldx, r, p, vp, ld, z, part_rz, part_rr);
    LAUNCH_CHECK();
    return true;
}

void launch_record_out(hipStream_t s, const RecMeta& m, const double* x, int64_t ldx, int k,
                       const double* GZ, const double* S0, const double* S1, int64_t ldn,
                       const double* theta_cols, double* y, int64_t ldy, double* part_dot,
                       const int* skip) {
    dim3 g(parts_for(m.nrec), k);
    k_record_out<<<g, kBlock, 0, s>>>(m, x, ldx, GZ, S0, S1, ldn, theta_cols, y, ldy, part_dot, skip);
    LAUNCH_CHECK();
}

void launch_precond(hipStream_t s, const RecMeta& m, const double* gdiag, const double* theta_cols,
                    const double* r, int64_t ldr, int k, double* z, int64_t ldz, double* part_rz,
                    double* part_rr, const PcgState* st, int tilde_v) {
    if (m.cap > 0) {
        dim3 g(precond_parts(m), k);
        k_precond_tiled<<<g, kBlock, sizeof(double) * (2 * m.cap + 2 * kTileBlocks), s>>>(m, gdiag, theta_cols, r, ldr, z,
                                                                       ldz, part_rz, part_rr, st, tilde_v);
    } else {
        dim3 g(parts_for(m.nblk), k);
        k_precond<<<g, kBlock, 0, s>>>(m.nblk, m.blk_off, m.blk_ind, m.times, gdiag, theta_cols, r,
                                       ldr, z, ldz, part_rz, part_rr, st, tilde_v);
    }
    LAUNCH_CHECK();
}

void launch_dot_parts(hipStream_t s, int64_t nrec, const double* a, int64_t lda, const double* b,
                      int64_t ldb, int k, double* part) {
    dim3 g(parts_for(nrec), k);
    k_dot_parts<<<g, kBlock, 0, s>>>(nrec, a, lda, b, ldb, part);
    LAUNCH_CHECK();
}

void launch_reduce_parts(hipStream_t s, const double* part, int nparts, int k, double* out) {
    if (k > 0) k_reduce_parts<<<k, kBlock, 0, s>>>(part, nparts, out, nullptr);
    LAUNCH_CHECK();
}

void launch_reduce_parts_skip(hipStream_t s, const double* part, int nparts, int k, double* out,
                              const int* skip) {
    if (k > 0) k_reduce_parts<<<k, kBlock, 0, s>>>(part, nparts, out, skip);
    LAUNCH_CHECK();
}

void launch_gram(hipStream_t s, int64_t nrec, const double* A, int64_t lda, int ka,
                 const double* B, int64_t ldb, int kb, double* part, double* out) {
    const int np = parts_for(nrec);
    dim3 g(np, ka * kb);
    k_gram_parts<<<g, kBlock, 0, s>>>(nrec, A, lda, B, ldb, kb, 0, part);
    LAUNCH_CHECK();
    if (ka * kb > 0) k_reduce_parts<<<ka * kb, kBlock, 0, s>>>(part, np, out, nullptr);
    LAUNCH_CHECK();
}

void launch_gram_grouped(hipStream_t s, int64_t nrec, const double* A, int64_t lda, int ka,
                         const double* B, int64_t ldb, int kb, double* part, double* out) {
    const int np = parts_for(nrec);
    dim3 g(np, ka * kb);
    k_gram_parts<<<g, kBlock, 0, s>>>(nrec, A, lda, B, ldb, kb, kb, part);
    LAUNCH_CHECK();
    if (ka * kb > 0) k_reduce_parts<<<ka * kb, kBlock, 0, s>>>(part, np, out, nullptr);
    LAUNCH_CHECK();
}

void launch_axpy_cols(hipStream_t s, int64_t nrec, double* y, int64_t ldy, const double* x,
                      int64_t ldx, int kx, const double* coef, int ky) {
    dim3 g(parts_for(nrec), ky);
    k_axpy_cols<<<g, kBlock, 0, s>>>(nrec, y, ldy, x, ldx, kx, coef, ky);
    LAUNCH_CHECK();
}

int launch_logdet_q(hipStream_t s, int32_t nblk, const int64_t* off, const int32_t* ind, const double* t,
                    const double* gdiag, const double* theta, double* part) {
    const int g = (int)std::min<int64_t>(std::max<int64_t>((nblk + kBlock - 1) / kBlock, 1), kMaxParts);
    k_logdet_q<<<g, kBlock, 0, s>>>(nblk, off, ind, t, gdiag, theta, part);
    HIP_CHECK(hipGetLastError());
    return g;
}

void launch_set_column(hipStream_t s, int j, int64_t off, int32_t ntiles, int64_t nrec, const double theta[5],
                       int64_t* col_off, int32_t* ntiles_col, int64_t* nrec_col, double* theta_cols) {
    k_set_column<<<1, 64, 0, s>>>(j, off, ntiles, nrec, theta[0], theta[1], theta[2], theta[3], theta[4], col_off,
                                  ntiles_col, nrec_col, theta_cols);
    LAUNCH_CHECK();
}

void launch_pcg_admit(hipStream_t s, PcgState* st, int j, const double* part_bb, int nparts_bb,
                      const double* part_rr, const double* part_rz, int nparts) {
    k_pcg_admit<<<1, 64, 0, s>>>(st, j, part_bb, nparts_bb, part_rr, part_rz, nparts);
    LAUNCH_CHECK();
}

void launch_pcg_init(hipStream_t s, PcgState* st, const double* part_bb, const double* part_rr,
                     const double* part_rz, int nparts, int k) {
    k_pcg_init<<<1, kPcgThreads, 0, s>>>(st, part_bb, part_rr, part_rz, nparts, k);
    LAUNCH_CHECK();
}


void launch_pcg_alpha(hipStream_t s, PcgState* st, const double* part_pvp, int nparts, int k) {
    k_pcg_alpha<<<1, kPcgThreads, 0, s>>>(st, part_pvp, nparts, k);
    LAUNCH_CHECK();
}

void launch_pcg_check(hipStream_t s, PcgState* st, const double* part_rr, int nparts, int k) {
    k_pcg_check<<<1, kPcgThreads, 0, s>>>(st, part_rr, nparts, k);
    LAUNCH_CHECK();
}

void launch_pcg_deactivate(hipStream_t s, PcgState* st, int j) {
    k_pcg_deactivate<<<1, 64, 0, s>>>(st, j);
    LAUNCH_CHECK();
}

void launch_pcg_check_beta(hipStream_t s, PcgState* st, const double* part_rr, int nparts_rr,
                           const double* part_rz, int nparts_rz, int k) {
    k_pcg_check_beta<<<1, kPcgThreads, 0, s>>>(st, part_rr, nparts_rr, part_rz, nparts_rz, k);
    LAUNCH_CHECK();
}

void launch_pcg_beta(hipStream_t s, PcgState* st, const double* part_rz, int nparts, int k) {
    k_pcg_beta<<<1, kPcgThreads, 0, s>>>(st, part_rz, nparts, k);
    LAUNCH_CHECK();
}

void launch_pcg_xr(hipStream_t s, PcgState* st, int64_t nrec, double* x, int64_t ldx, double* r,
                   const double* p, const double* vp, int64_t ld, int k, double* part_rr,
                   const int64_t* col_off, const int64_t* col_cnt) {
    dim3 g(parts_for(nrec), k);
    k_pcg_xr<<<g, kBlock, 0, s>>>(st, nrec, x, ldx, r, p, vp, ld, part_rr, col_off, col_cnt);
    LAUNCH_CHECK();
}



void launch_pcg_p(hipStream_t s, PcgState* st, int64_t nrec, double* p, const double* z,
                  int64_t ld, int k, const int64_t* col_off, const int64_t* col_cnt) {
    dim3 g(parts_for(nrec), k);
    k_pcg_p<<<g, kBlock, 0, s>>>(st, nrec, p, z, ld, col_off, col_cnt);
    LAUNCH_CHECK();
}

void launch_bhat(hipStream_t s, int32_t nblk, const int64_t* blk_off, const int32_t* blk_ind,
                 const double* times, const double* v, const double* theta, double* bhat) {
    k_bhat<<<parts_for(nblk), kBlock, 0, s>>>(nblk, blk_off, blk_ind, times, v, theta, bhat);
    LAUNCH_CHECK();
}

void launch_f_theta(hipStream_t s, int32_t nblk, const int64_t* blk_off, const int32_t* blk_ind,
                    const double* times, const double* py, int64_t ldpy, const double* uhat,
                    int64_t ldu, int ncols, double inv_sg, double* f, int64_t ldf) {
    dim3 g(parts_for(nblk), ncols);
    k_f_theta<<<g, kBlock, 0, s>>>(nblk, blk_off, blk_ind, times, py, ldpy, uhat, ldu, inv_sg, f,
                                   ldf);
    LAUNCH_CHECK();
}

void launch_copy_cols(hipStream_t s, int64_t rows, const double* src, int64_t lds, double* dst,
                      int64_t ldd, int k) {
    dim3 g(parts_for(rows), k);
    k_copy_cols<<<g, kBlock, 0, s>>>(rows, src, lds, dst, ldd);
    LAUNCH_CHECK();
}

void launch_sub_cols(hipStream_t s, int64_t rows, const double* a, int64_t lda, const double* b,
                     int64_t ldb, double* d, int64_t ldd, int k) {
    dim3 g(parts_for(rows), k);
    k_sub_cols<<<g, kBlock, 0, s>>>(rows, a, lda, b, ldb, d, ldd);
    LAUNCH_CHECK();
}

void launch_fill_cols(hipStream_t s, int64_t rows, double* d, int64_t ldd, int k, double v) {
    dim3 g(parts_for(rows), k);
    k_fill_cols<<<g, kBlock, 0, s>>>(rows, d, ldd, v);
    LAUNCH_CHECK();
}

void launch_copy16(hipStream_t s, void* dst, const void* src, int64_t bytes) {
    const int64_t n16 = bytes / 16;
    k_copy16<<<2048, kBlock, 0, s>>>(static_cast<uint4*>(dst), static_cast<const uint4*>(src), n16);
    LAUNCH_CHECK();
}

void launch_fill(hipStream_t s, double* dst, int64_t count, double v) {
    k_fill<<<parts_for(count), kBlock, 0, s>>>(dst, count, v);
    LAUNCH_CHECK();
}

void launch_zero_slots(hipStream_t s, double* a, double* b, double* c, int k, int from, int to) {
    if (to <= from || k <= 0) return;
    k_zero_slots<<<dim3(1, 3), kBlock, 0, s>>>(a, b, c, k, from, to);
    LAUNCH_CHECK();
}

}  // namespace globe
