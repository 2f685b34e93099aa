// Monte Carlo probes of the AI-REML trace estimator, generated on the device.
//
// Random bits come from Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11): a counter-based generator, so every (probe,
// record) value is a pure function of the seed and its counter, independent of launch
// shape. Counter words: (index low, index high, probe, stream), key = the 64-bit seed.
//   Rademacher (kind 1): z[s][r] = +1 if bit 31 of word 0 of (r, s, 0) is set, else -1.
//   Gaussian tilde-V (kind 0, the reference's sample_Yvs, varCompEst.cpp:775-797):
//     per individual i, (i, s, 1) gives z1 = g(w0, w1), z2a = g'(w0, w1), z2b = g(w2, w3);
//     per record r, (r, s, 2) gives z3 = g(w0, w1); g / g' are the Box-Muller pair
//     sqrt(-2 ln u) cos(2 pi v) / sin(2 pi v) with u = (w + 1/2) 2^-32;
//     y_r = sqrt(sig2g) z1 + sqrt(sig2e) z3 + z2a (l00 + l10 t_r) + z2b l11 t_r, with
//     [[l00, 0], [l10, l11]] the Cholesky factor of B (compute_chol_2x2, :800-813; a
//     negative l11^2 is clipped to 0, and sig2b0 = 0 drops the slope terms).
//   exact (kind 2): y[s][r] = sqrt(nrec) for r == s, else 0 (nmc == nrec).
#include <cmath>

#include "globe_internal.h"

namespace globe {

namespace {

struct U4 {
    uint32_t w[4];
};

__device__ inline U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                   uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{{c0, c1, c2, c3}};
}

__device__ inline double unit_open(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

__global__ __launch_bounds__(256) void k_probes(int kind, int64_t nrec, uint64_t seed,
                                                const int32_t* __restrict__ indiv,
                                                const double* __restrict__ times, double sg, double se,
                                                double l00, double l10, double l11,
                                                double* __restrict__ Y, int64_t ldy) {
    const int s = blockIdx.y;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int64_t r = blockIdx.x * 256ll + threadIdx.x; r < nrec; r += (int64_t)gridDim.x * 256) {
        double y;
        if (kind == 1) {
            const U4 o = philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), (uint32_t)s, 0u, k0, k1);
            y = (o.w[0] >> 31) ? 1.0 : -1.0;
        } else if (kind == 2) {
            y = r == s ? sqrt((double)nrec) : 0.0;
        } else {
            const int64_t i = indiv[r];
            const U4 a = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), (uint32_t)s, 1u, k0, k1);
            const U4 b = philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), (uint32_t)s, 2u, k0, k1);
            const double two_pi = 6.283185307179586476925286766559;
            const double ra = sqrt(-2.0 * log(unit_open(a.w[0])));
            const double z1 = ra * cos(two_pi * unit_open(a.w[1]));
            const double z2a = ra * sin(two_pi * unit_open(a.w[1]));
            const double z2b = sqrt(-2.0 * log(unit_open(a.w[2]))) * cos(two_pi * unit_open(a.w[3]));
            const double z3 = sqrt(-2.0 * log(unit_open(b.w[0]))) * cos(two_pi * unit_open(b.w[1]));
            const double t = times[r];
            y = sqrt(sg) * z1 + sqrt(se) * z3 + z2a * (l00 + l10 * t) + z2b * l11 * t;
        }
        Y[s * ldy + r] = y;
    }
}

// Probes b = F u of the log-determinant estimator, F F^T = Q (the block preconditioner:
// Q_i = sig2e I + U_i C_i U_i^T, U_i = [1, t], C_i = [[sig2g G_ii + sig2b0, sigb01],
// [sigb01, sig2b1]]): b_r = sqrt(sig2e) u3_r + ua_i (k00 + k10 t_r) + ub_i k11 t_r with
// [[k00, 0], [k10, k11]] the Cholesky factor of C_i (a negative pivot clipped to 0).
// Probe s has the global index g = first + s. kind 1: u3 = sign of bit 31 of word 0 of
// (r, g, 3), ua / ub = signs of words 0 / 1 of (i, g, 4); kind 0: Box-Muller normals on
// the same counters; kind 2: column g of F (g < nrec: sqrt(sig2e) e_g; else, for
// g = nrec + 2 i + a, the column a of U_i K_i on the records of individual i).
__global__ __launch_bounds__(256) void k_logdet_probes(int kind, int64_t nrec, int64_t first, uint64_t seed,
                                                       const int32_t* __restrict__ indiv,
                                                       const double* __restrict__ times,
                                                       const double* __restrict__ gdiag, double se,
                                                       double sg, double sb0, double sb01, double sb1,
                                                       double* __restrict__ Y, int64_t ldy) {
    const int s = blockIdx.y;
    const int64_t g = first + s;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t r = blockIdx.x * 256ll + threadIdx.x; r < nrec; r += (int64_t)gridDim.x * 256) {
        const int64_t i = indiv[r];
        const double t = times[r];
        const double c00 = sg * gdiag[i] + sb0;
        const double k00 = c00 > 0.0 ? sqrt(c00) : 0.0;
        const double k10 = k00 > 0.0 ? sb01 / k00 : 0.0;
        const double d11 = sb1 - k10 * k10;
        const double k11 = d11 > 0.0 ? sqrt(d11) : 0.0;
        double y;
        if (kind == 2) {
            if (g < nrec) {
                y = r == g ? sqrt(se) : 0.0;
            } else {
                const int64_t j = g - nrec;
                y = i != j / 2 ? 0.0 : ((j & 1) == 0 ? k00 + k10 * t : k11 * t);
            }
        } else {
            const U4 w = philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), (uint32_t)g, 3u, k0, k1);
            const U4 v = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), (uint32_t)g, 4u, k0, k1);
            double u3, ua, ub;
            if (kind == 1) {
                u3 = (w.w[0] >> 31) ? 1.0 : -1.0;
                ua = (v.w[0] >> 31) ? 1.0 : -1.0;
                ub = (v.w[1] >> 31) ? 1.0 : -1.0;
            } else {
                u3 = sqrt(-2.0 * log(unit_open(w.w[0]))) * cos(two_pi * unit_open(w.w[1]));
                const double ra = sqrt(-2.0 * log(unit_open(v.w[0])));
                ua = ra * cos(two_pi * unit_open(v.w[1]));
                ub = ra * sin(two_pi * unit_open(v.w[1]));
            }
            y = sqrt(se) * u3 + ua * (k00 + k10 * t) + ub * k11 * t;
        }
        Y[s * ldy + r] = y;
    }
}

}  // namespace

void launch_logdet_probes(hipStream_t s, int kind, int64_t nrec, int64_t first, int count, uint64_t seed,
                          const int32_t* indiv, const double* times, const double* gdiag,
                          const double theta[5], double* Y, int64_t ldy) {
    if (nrec <= 0 || count <= 0) return;
    const unsigned gx = (unsigned)std::min<int64_t>((nrec + 255) / 256, 1024);
    k_logdet_probes<<<dim3(gx, count), 256, 0, s>>>(kind, nrec, first, seed, indiv, times, gdiag, theta[0],
                                                    theta[1], theta[2], theta[3], theta[4], Y, ldy);
    HIP_CHECK(hipGetLastError());
}

void launch_probes(hipStream_t s, int kind, int64_t nrec, int nmc, uint64_t seed, const int32_t* indiv,
                   const double* times, const double theta[5], double* Y, int64_t ldy) {
    if (nrec <= 0 || nmc <= 0) return;
    const double se = theta[0], sg = theta[1], sb0 = theta[2], sb01 = theta[3], sb1 = theta[4];
    const double l00 = sb0 > 0.0 ? std::sqrt(sb0) : 0.0;
    const double l10 = sb0 > 0.0 ? sb01 / l00 : 0.0;
    double d11 = sb1 - (sb0 > 0.0 ? sb01 * sb01 / sb0 : 0.0);
    if (d11 < 0.0) d11 = 0.0;
    const unsigned gx = (unsigned)std::min<int64_t>((nrec + 255) / 256, 1024);
    k_probes<<<dim3(gx, nmc), 256, 0, s>>>(kind, nrec, seed, indiv, times, sg, se, l00, l10, std::sqrt(d11),
                                           Y, ldy);
    HIP_CHECK(hipGetLastError());
}

}  // namespace globe
