// Near-field (direct) interactions of the box FGT on the GPU (FP64).
//
// For each target leaf at or above the cutoff level, every source leaf of
// its modified list 1 (plus periodic images of level-1 sources) contributes
// pot(k1,k2) += sum_j2 Ty(j2,k2) sum_j1 Tx(j1,k1) f(j1,j2), with the 1-D
// tables of the target level picked by the geometric offsets (ix, iy)
// (bfgt/boxfgt.f:923-1071, bfgt/bfgt_local.f:315-545).  The tables are the
// eps-masked dense tables built on the device, so the dense contraction sums
// the same terms as the reference's sparse loops.  One block owns one target
// and accumulates all its sources in registers before a single
// read-modify-write of pot/grad/hess: no atomics.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bfgt_internal.h"

namespace bfgt {

namespace {

constexpr int kNearThreads = 256;
constexpr int kNearMaxQ = (kMaxOrder * kMaxOrder + kNearThreads - 1) / kNearThreads;

__global__ __launch_bounds__(kNearThreads) void k_near(int n, int nd, int ntab, const int* __restrict__ tgt,
                                                       const int* __restrict__ tlev,
                                                       const int* __restrict__ ptr,
                                                       const NearPair* __restrict__ pairs,
                                                       const double* __restrict__ tloc,
                                                       const double* __restrict__ fvals, double* __restrict__ pot,
                                                       double* __restrict__ grad, double* __restrict__ hess,
                                                       double oscale, int half, const int* __restrict__ tfresh,
                                                       int accum) {
    extern __shared__ double smem[];
    const int nn = n * n;
    double* sf = smem;       // nn
    double* sff = sf + nn;   // ntab * nn : [t][k1 + n*j2]
    const int b = tgt[blockIdx.x];
    const int lev = tlev[blockIdx.x];
    // an overwriting execution writes the outputs of a target that no PW
    // evaluation wrote, instead of adding to them
    const bool fresh = accum == 0 && tfresh[blockIdx.x] != 0;
    const int p0 = ptr[blockIdx.x], p1 = ptr[blockIdx.x + 1];
    const int tid = threadIdx.x;
    const double* T = tloc + (size_t)lev * (2 * half + 1) * 3 * nn;
    for (int d = 0; d < nd; d++) {
        double acc[kNearMaxQ][6];
#pragma unroll
        for (int r = 0; r < kNearMaxQ; r++)
#pragma unroll
            for (int c = 0; c < 6; c++) acc[r][c] = 0.0;
        for (int p = p0; p < p1; p++) {
            const NearPair pr = pairs[p];
            const double* fs = fvals + (size_t)pr.src * nn * nd;
            __syncthreads();
            for (int e = tid; e < nn; e += kNearThreads) sf[e] = fs[(size_t)e * nd + d];
            __syncthreads();
            const double* Tx = T + (size_t)(pr.ix + half) * 3 * nn;
            const double* Ty = T + (size_t)(pr.iy + half) * 3 * nn;
            // x pass: ff_t(k1, j2) = sum_j1 Tx_t(k1, j1) f(j1, j2)
            for (int e = tid; e < ntab * nn; e += kNearThreads) {
                int t = e / nn, r = e % nn;
                int k1 = r % n, j2 = r / n;
                const double* row = Tx + (size_t)t * nn + k1 * n;
                double s = 0.0;
                for (int j1 = 0; j1 < n; j1++) s = fma(row[j1], sf[j1 + n * j2], s);
                sff[e] = s;
            }
            __syncthreads();
            // y pass into the per-point accumulators
#pragma unroll
            for (int r = 0; r < kNearMaxQ; r++) {
                int q = tid + r * kNearThreads;
                if (q < nn) {
                    int k1 = q % n, k2 = q / n;
                    const double* y0 = Ty + k2 * n;
                    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
                    for (int j2 = 0; j2 < n; j2++) {
                        double f0 = sff[k1 + n * j2];
                        double t0 = y0[j2];
                        s0 = fma(t0, f0, s0);
                        if (ntab >= 2) {
                            double f1 = sff[nn + k1 + n * j2];
                            double t1 = y0[nn + j2];
                            s1 = fma(t0, f1, s1);
                            s2 = fma(t1, f0, s2);
                            if (ntab >= 3) {
                                double f2 = sff[2 * nn + k1 + n * j2];
                                double t2 = y0[2 * nn + j2];
                                s3 = fma(t0, f2, s3);
                                s4 = fma(t1, f1, s4);
                                s5 = fma(t2, f0, s5);
                            }
                        }
                    }
                    acc[r][0] += s0;
                    acc[r][1] += s1;
                    acc[r][2] += s2;
                    acc[r][3] += s3;
                    acc[r][4] += s4;
                    acc[r][5] += s5;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kNearMaxQ; r++) {
            int q = tid + r * kNearThreads;
            if (q < nn && p1 > p0) {
                size_t pb = (size_t)b * nn + q;
                auto put = [&](double& v, double x) { v = (fresh ? 0.0 : v) + oscale * x; };
                put(pot[pb * nd + d], acc[r][0]);
                if (ntab >= 2) {
                    put(grad[(pb * 2 + 0) * nd + d], acc[r][1]);
                    put(grad[(pb * 2 + 1) * nd + d], acc[r][2]);
                }
                if (ntab >= 3) {
                    put(hess[(pb * 3 + 0) * nd + d], acc[r][3]);
                    put(hess[(pb * 3 + 1) * nd + d], acc[r][4]);
                    put(hess[(pb * 3 + 2) * nd + d], acc[r][5]);
                }
            }
        }
    }
}

// Adjoint of the near field: one block per source leaf accumulates, over
// every (target, offset) pair that reads it,
//   fbar(j1, j2) += sum_o sum_{k1,k2} Tx_{x(o)}(k1, j1) Ty_{y(o)}(k2, j2) g_o(k1, k2)
// with the target's tables read transposed and o over the channels pot (0,0),
// gx (1,0), gy (0,1), hxx (2,0), hxy (1,1), hyy (0,2) (x table, y table).
__global__ __launch_bounds__(kNearThreads) void k_near_adj(int n, int nd, int ntab, const int* __restrict__ src,
                                                           const int* __restrict__ ptr,
                                                           const NearPairT* __restrict__ pairs,
                                                           const double* __restrict__ tloc,
                                                           const double* __restrict__ gpot,
                                                           const double* __restrict__ ggrad,
                                                           const double* __restrict__ ghess,
                                                           double* __restrict__ fbar) {
    extern __shared__ double smem[];
    const int nn = n * n;
    const int nout = ntab == 1 ? 1 : (ntab == 2 ? 3 : 6);
    double* sg = smem;            // nout * nn : [o][k1 + n*k2]
    double* sz = sg + nout * nn;  // ntab * nn : [u][j1 + n*k2]
    const int s = src[blockIdx.x];
    const int p0 = ptr[blockIdx.x], p1 = ptr[blockIdx.x + 1];
    const int tid = threadIdx.x;
    for (int d = 0; d < nd; d++) {
        double acc[kNearMaxQ];
#pragma unroll
        for (int r = 0; r < kNearMaxQ; r++) acc[r] = 0.0;
        for (int p = p0; p < p1; p++) {
            const NearPairT pr = pairs[p];
            const double* T = tloc + (size_t)pr.lev * 13 * 3 * nn;
            const double* Tx = T + (size_t)(pr.ix + 6) * 3 * nn;
            const double* Ty = T + (size_t)(pr.iy + 6) * 3 * nn;
            const size_t tb = (size_t)pr.tgt * nn;
            __syncthreads();
            for (int e = tid; e < nn; e += kNearThreads) {
                size_t pb = tb + e;
                sg[e] = gpot[pb * nd + d];
                if (ntab >= 2) {
                    sg[nn + e] = ggrad[(pb * 2 + 0) * nd + d];
                    sg[2 * nn + e] = ggrad[(pb * 2 + 1) * nd + d];
                }
                if (ntab >= 3) {
                    sg[3 * nn + e] = ghess[(pb * 3 + 0) * nd + d];
                    sg[4 * nn + e] = ghess[(pb * 3 + 1) * nd + d];
                    sg[5 * nn + e] = ghess[(pb * 3 + 2) * nd + d];
                }
            }
            __syncthreads();
            // x pass: Z_u(j1, k2) = sum over channels with y table u of Tx_x(o)^T g_o
            for (int e = tid; e < ntab * nn; e += kNearThreads) {
                int u = e / nn, r = e % nn;
                int j1 = r % n, k2 = r / n;
                double z = 0.0;
                for (int k1 = 0; k1 < n; k1++) {
                    int g = k1 + n * k2;
                    int ti = k1 * n + j1;
                    if (u == 0) {
                        z = fma(Tx[ti], sg[g], z);
                        if (ntab >= 2) z = fma(Tx[nn + ti], sg[nn + g], z);
                        if (ntab >= 3) z = fma(Tx[2 * nn + ti], sg[3 * nn + g], z);
                    } else if (u == 1) {
                        z = fma(Tx[ti], sg[2 * nn + g], z);
                        if (ntab >= 3) z = fma(Tx[nn + ti], sg[4 * nn + g], z);
                    } else {
                        z = fma(Tx[ti], sg[5 * nn + g], z);
                    }
                }
                sz[e] = z;
            }
            __syncthreads();
            // y pass: fbar(j1, j2) += sum_u Ty_u^T Z_u
#pragma unroll
            for (int r = 0; r < kNearMaxQ; r++) {
                int q = tid + r * kNearThreads;
                if (q < nn) {
                    int j1 = q % n, j2 = q / n;
                    double a = 0.0;
                    for (int u = 0; u < ntab; u++)
                        for (int k2 = 0; k2 < n; k2++) a = fma(Ty[u * nn + k2 * n + j2], sz[u * nn + j1 + n * k2], a);
                    acc[r] += a;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kNearMaxQ; r++) {
            int q = tid + r * kNearThreads;
            if (q < nn && p1 > p0) fbar[((size_t)s * nn + q) * nd + d] += acc[r];
        }
    }
}

// ---------------------------------------------------------------------------
// matrix-core near field (n <= 16)
// ---------------------------------------------------------------------------
//
// The tables are read from zero-padded copies, tnear[level][13][3][16][16]
// ([k+6][t][target k][source j]) for the adjoint and its transpose tnearT
// ([k+6][t][source j][target k]) for the forward, so that every table load
// of 64 lanes covers four 16-entry rows; out-of-range rows and columns add
// exact zeros; leaf values are read at clamped indices (finite values that
// only meet those zeros).  v_mfma_f64_16x16x4f64 lane layout as in the PW
// kernels: A holds A[c][q], B holds B[q][c], element r of the accumulator
// holds D[4r + q][c] (lane l, q = l/16, c = l%16).

typedef double v4d __attribute__((ext_vector_type(4)));
__device__ inline v4d mfma(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
constexpr int near_chan(int ntab) { return ntab == 1 ? 1 : (ntab == 2 ? 3 : 6); }
constexpr int near_x(int o) { return o == 1 || o == 4 ? 1 : (o == 3 ? 2 : 0); }
constexpr int near_y(int o) { return o == 2 || o == 4 ? 1 : (o == 5 ? 2 : 0); }
constexpr int kNearWaves = 4;
// The part of the x tables a wave of a split target handles: 0 = table 0,
// 1 = tables 1 and 2, -1 = all (no split).
constexpr bool near_part_has(int tw, int t) { return tw < 0 || (tw == 0 ? t == 0 : t >= 1); }

// One wave per (target, density).  For each (source, offsets) pair:
//   FF_t^T = F^T Tx_t^T           (x contraction, K = source x index)
//   out_o += Ty_{y(o)} FF_{x(o)}   (y contraction, K = source y index)
// where element s of FF_t^T's accumulator is the B operand of K-step s of
// the second product.  Channels o: pot (0,0), gx (1,0), gy (0,1), hxx (2,0),
// hxy (1,1), hyy (0,2) as (x table, y table).  NKS = ceil(n/4).
template <int NKS, int NTAB, int NSPLIT>
__global__ __launch_bounds__(64 * kNearWaves) void k_near_mfma(
    int n, int nd, int nitems, const int* __restrict__ tgt, const int* __restrict__ tlev,
    const int* __restrict__ ptr, const NearPair* __restrict__ pairs, const double* __restrict__ tnearT,
    const double* __restrict__ fvals, double* __restrict__ pot, double* __restrict__ grad,
    double* __restrict__ hess, double oscale, const int* __restrict__ tfresh, int accum) {
    constexpr int NOUT = near_chan(NTAB);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    // wave-uniform: schedule entries are then read with scalar loads
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kNearWaves + wave);
    if (item >= nitems) return;  // no block-wide barriers in this kernel
    const int tw = item % NSPLIT, it = item / NSPLIT / nd, d = item / NSPLIT % nd;
    const int q = lane >> 4, c = lane & 15;
    const int b = tgt[it];
    const bool fresh = accum == 0 && tfresh[it] != 0;  // as in k_near
    const int p0 = ptr[it], p1 = ptr[it + 1];
    const double* T = tnearT + (size_t)tlev[it] * 13 * 3 * 256;
    const int nn = n * n;
    const int cl = min(c, n - 1);
    // TW: this wave's part of the x tables (near_part_has): their x
    // contractions and the outputs o whose x table near_x(o) is one of them
    auto body = [&](auto twc) {
        constexpr int TW = decltype(twc)::value;
        v4d out[NOUT];
#pragma unroll
        for (int o = 0; o < NOUT; o++) out[o] = v4d{0, 0, 0, 0};
        // two pairs per iteration so that their loads overlap (a repeated last
        // pair of an odd count contributes with weight zero)
        for (int p = p0; p < p1; p += 2) {
            const bool two = p + 1 < p1;
            const NearPair pr[2] = {pairs[p], pairs[two ? p + 1 : p]};
            v4d FF[2][NTAB];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double* fs = fvals + (size_t)pr[h].src * nn * nd + d;
                const double* Tx = T + (size_t)(pr[h].ix + 6) * 3 * 256;
                const double wgt = (h == 0 || two) ? 1.0 : 0.0;
#pragma unroll
                for (int t = 0; t < NTAB; t++) FF[h][t] = v4d{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < NKS; s++) {
                    const int j1 = 4 * s + q;
                    // A = F^T: A[j2 = c][j1]; B = Tx_t^T: B[j1][k1 = c] = Tx_t[c][j1] (transposed table)
                    const double a = wgt * fs[(size_t)(min(j1, n - 1) + n * cl) * nd];
#pragma unroll
                    for (int t = 0; t < NTAB; t++)
                        if (near_part_has(TW, t)) FF[h][t] = mfma(a, Tx[t * 256 + j1 * 16 + c], FF[h][t]);
                }
            }
            // FF_t^T: element s of lane (q, c) = FF_t^T[j2 = 4s + q][k1 = c]
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double* Ty = T + (size_t)(pr[h].iy + 6) * 3 * 256;
#pragma unroll
                for (int s = 0; s < NKS; s++) {
                    const int j2 = 4 * s + q;
                    double ty[NTAB];
                    // A[k2 = c][j2] = Ty[c][j2]; x tables 1 and 2 meet only the y
                    // tables u with 1 + u < NTAB
#pragma unroll
                    for (int u = 0; u < NTAB; u++)
                        if (TW <= 0 || u + 1 < NTAB) ty[u] = Ty[u * 256 + j2 * 16 + c];
#pragma unroll
                    for (int o = 0; o < NOUT; o++)
                        if (near_part_has(TW, near_x(o)))
                            out[o] = mfma(ty[near_y(o)], FF[h][near_x(o)][s], out[o]);
                }
            }
        }
        // element r of lane (q, c) = out[k1 = c][k2 = 4r + q].  The outputs are
        // read at clamped points for all r first (one memory round trip), then
        // added and stored where valid.
        if (p1 > p0) {
            double* dst[NOUT];
            dst[0] = pot;
            if constexpr (NTAB >= 2) dst[1] = dst[2] = grad;
            if constexpr (NTAB >= 3) dst[3] = dst[4] = dst[5] = hess;
            constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
            double old[4][NOUT];
            if (!fresh) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const size_t pb = (size_t)b * nn + min(c, n - 1) + (size_t)n * min(4 * r + q, n - 1);
#pragma unroll
                    for (int o = 0; o < NOUT; o++)
                        if (near_part_has(TW, near_x(o)))
                            old[r][o] = dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int o = 0; o < NOUT; o++) old[r][o] = 0.0;
            }
            if (c < n) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int k2 = 4 * r + q;
                    if (k2 < n) {
                        const size_t pb = (size_t)b * nn + c + (size_t)n * k2;
#pragma unroll
                        for (int o = 0; o < NOUT; o++)
                            if (near_part_has(TW, near_x(o)))
                                dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d] = old[r][o] + oscale * out[o][r];
                    }
                }
            }
        }
    };
    if constexpr (NSPLIT == 1) {
        body(std::integral_constant<int, -1>{});
    } else {
        static_assert(NSPLIT == 2 && NTAB >= 2, "x table 0 | the others");
        if (tw == 0) body(std::integral_constant<int, 0>{});
        else body(std::integral_constant<int, 1>{});
    }
}

// Split-FP32 near field (BFGT_PREC_SPLIT32, n <= 16): one wave per
// (target, density), the products of k_near_mfma on the FP32 matrix cores
// with hi/lo operands (three products per table product, two where the
// second factor is the FP32 accumulator of the first), accumulated in FP32
// per target and added to the FP64 outputs.  v_mfma_f32_16x16x4f32 holds
// accumulator element r of lane (q, c) at row 4q + r: the x contraction takes
// source x index j1 = 4s + q in K-step s, the y contraction source y index
// j2 = 4q + r in K-step r (all four: rows past n meet zero table columns).
template <int NKS, int NTAB>
__global__ __launch_bounds__(64 * kNearWaves) void k_near_split(
    int n, int nd, int nitems, const int* __restrict__ tgt, const int* __restrict__ tlev,
    const int* __restrict__ ptr, const NearPair* __restrict__ pairs, const double* __restrict__ tnearT,
    const double* __restrict__ fvals, double* __restrict__ pot, double* __restrict__ grad,
    double* __restrict__ hess, double oscale, const int* __restrict__ tfresh, int accum) {
    constexpr int NOUT = near_chan(NTAB);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kNearWaves + wave);
    if (item >= nitems) return;  // no block-wide barriers in this kernel
    const int it = item / nd, d = item % nd;
    const int q = lane >> 4, c = lane & 15;
    const int b = tgt[it];
    const bool fresh = accum == 0 && tfresh[it] != 0;  // as in k_near
    const int p0 = ptr[it], p1 = ptr[it + 1];
    const double* T = tnearT + (size_t)tlev[it] * 13 * 3 * 256;
    const int nn = n * n;
    const int cl = min(c, n - 1);
    v4f out[NOUT];
#pragma unroll
    for (int o = 0; o < NOUT; o++) out[o] = v4f{0, 0, 0, 0};
    for (int p = p0; p < p1; p++) {
        const NearPair pr = pairs[p];
        const double* fs = fvals + (size_t)pr.src * nn * nd + d;
        const double* Tx = T + (size_t)(pr.ix + 6) * 3 * 256;
        const double* Ty = T + (size_t)(pr.iy + 6) * 3 * 256;
        // FF_t^T[j2 = row][k1 = col] = sum_j1 F[j1][j2] Tx_t[k1][j1]
        v4f FF[NTAB];
#pragma unroll
        for (int t = 0; t < NTAB; t++) FF[t] = v4f{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < NKS; s++) {
            const int j1 = 4 * s + q;
            const Split a = split(fs[(size_t)(min(j1, n - 1) + n * cl) * nd]);
#pragma unroll
            for (int t = 0; t < NTAB; t++) FF[t] = mfma3(a, split(Tx[t * 256 + j1 * 16 + c]), FF[t]);
        }
        // element r of FF_t on lane (q, c) = FF_t^T[j2 = 4q + r][k1 = c]
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int j2 = drow32(q, r);
            Split ty[NTAB];
#pragma unroll
            for (int u = 0; u < NTAB; u++) ty[u] = split(Ty[u * 256 + j2 * 16 + c]);
#pragma unroll
            for (int o = 0; o < NOUT; o++) out[o] = mfma2(ty[near_y(o)], FF[near_x(o)][r], out[o]);
        }
    }
    // element r of out_o on lane (q, c): target point k1 = c, k2 = 4q + r
    if (p1 > p0 && c < n) {
        double* dst[NOUT];
        dst[0] = pot;
        if constexpr (NTAB >= 2) dst[1] = dst[2] = grad;
        if constexpr (NTAB >= 3) dst[3] = dst[4] = dst[5] = hess;
        constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int k2 = drow32(q, r);
            if (k2 < n) {
                const size_t pb = (size_t)b * nn + c + (size_t)n * k2;
#pragma unroll
                for (int o = 0; o < NOUT; o++) {
                    double& v = dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d];
                    v = (fresh ? 0.0 : v) + oscale * (double)out[o][r];
                }
            }
        }
    }
}

// n <= 8 and two or more densities: one wave per (target, pair of densities
// d0, d0 + 1, tile).  At n <= 8 a 16x16 tile of one density and one table is a
// quarter full, so the x contraction holds the two densities in its row
// halves (h) and two x tables in its column halves (g):
//   FF_X[(h, j2)][(g, k1)] = sum_j1 F_{d0+h}[j1][j2] Tx_{tx(X,g)}[k1][j1]
// and the y contraction of density h takes the K-steps of its row half and
// holds two y tables in its row halves (v):
//   Q_{h,X}[(v, k2)][(g, k1)] = sum_j2 Ty_{ty(X,v)}[k2][j2] FF_X[(h, j2)][(g, k1)]
// Tiles X0: x tables (0 | 1), y tables (0 | 1): pot, gx / gy, hxy;
//       X1: x tables (2 | 0), y tables (0 | 2): hxx, - / -, hyy.
// Three tables take 4 + 8 products for two densities where one density per
// wave and one table per tile take 2 x 18.  Quadrants marked "-", those of
// a missing density d0 + 1 = nd and, for pot only, all but (0, 0) hold
// finite values that are never stored.
template <int NTAB>
__device__ constexpr int pack_tx(int X, int g) { return X == 0 ? (NTAB == 1 ? 0 : g) : (g ? 0 : 2); }
template <int NTAB>
__device__ constexpr int pack_ty(int X, int v) { return X == 0 ? (NTAB == 1 ? 0 : v) : (v ? 2 : 0); }
// output channel of quadrant (v, g) of tile X: 0 pot, 1 gx, 2 gy, 3 hxx, 4 hxy, 5 hyy; -1 none
template <int NTAB>
__device__ constexpr int pack_chan(int X, int g, int v) {
    if (X == 1) return g != v ? -1 : (g == 0 ? 3 : 5);
    if (NTAB == 1) return g == 0 && v == 0 ? 0 : -1;
    return g + v >= NTAB ? -1 : (g + v) * (g + v + 1) / 2 + v;
}

template <int NKS, int NTAB>
__global__ __launch_bounds__(64 * kNearWaves) void k_near_pack(
    int n, int nd, int nitems, const int* __restrict__ tgt, const int* __restrict__ tlev,
    const int* __restrict__ ptr, const NearPair* __restrict__ pairs, const double* __restrict__ tnearT,
    const double* __restrict__ fvals, double* __restrict__ pot, double* __restrict__ grad,
    double* __restrict__ hess, double oscale, const int* __restrict__ tfresh, int accum) {
    static_assert(NKS <= 2, "two densities per tile need n <= 8");
    constexpr int NX = NTAB == 3 ? 2 : 1;  // tiles, one wave each
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kNearWaves + wave);
    if (item >= nitems) return;  // no block-wide barriers in this kernel
    const int ndp = (nd + 1) / 2;
    const int X = item % NX, it = item / NX / ndp, d0 = 2 * (item / NX % ndp);
    const int q = lane >> 4, c = lane & 15, h = c >> 3, cc = c & 7;
    const int b = tgt[it];
    const bool fresh = accum == 0 && tfresh[it] != 0;  // as in k_near
    const int p0 = ptr[it], p1 = ptr[it + 1];
    const double* T = tnearT + (size_t)tlev[it] * 13 * 3 * 256;
    const int nn = n * n;
    const int cl = min(cc, n - 1);
    // A operand of the x contraction: row c = (density h, source y index cc)
    const int dh = min(d0 + h, nd - 1);
    // per lane: x table of column half h, y table of row half h
    const int tx = h ? pack_tx<NTAB>(X, 1) : pack_tx<NTAB>(X, 0);
    const int ty = h ? pack_ty<NTAB>(X, 1) : pack_ty<NTAB>(X, 0);
    v4d out[2] = {v4d{0, 0, 0, 0}, v4d{0, 0, 0, 0}};
    for (int p = p0; p < p1; p += 2) {
        const bool two = p + 1 < p1;
        const NearPair pr[2] = {pairs[p], pairs[two ? p + 1 : p]};
        v4d FF[2];
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const double* fs = fvals + (size_t)pr[s2].src * nn * nd + dh;
            const double* Tx = T + (size_t)(pr[s2].ix + 6) * 3 * 256;
            const double wgt = (s2 == 0 || two) ? 1.0 : 0.0;
            FF[s2] = v4d{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < NKS; s++) {
                const int j1 = 4 * s + q;
                const double a = wgt * fs[(size_t)(min(j1, n - 1) + n * cl) * nd];
                FF[s2] = mfma(a, Tx[tx * 256 + j1 * 16 + cc], FF[s2]);
            }
        }
        // element s of FF on lane (q, c) = FF[4s + q][c]: rows of density
        // half s / 2, source y indices 4 (s % 2) + q
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const double* Ty = T + (size_t)(pr[s2].iy + 6) * 3 * 256;
#pragma unroll
            for (int ss = 0; ss < NKS; ss++) {
                const int j2 = 4 * ss + q;
                const double ya = Ty[ty * 256 + j2 * 16 + cc];  // A[(v, k2 = cc)][j2]
#pragma unroll
                for (int hh = 0; hh < 2; hh++) out[hh] = mfma(ya, FF[s2][2 * hh + ss], out[hh]);
            }
        }
    }
    // element r of Q_{hh,X} on lane (q, c): density d0 + hh, y-table half
    // v = r / 2, target y index k2 = 4 (r % 2) + q, x-table half h, target x
    // index k1 = cc
    if (p1 > p0) {
        int o[4];
#pragma unroll
        for (int r = 0; r < 4; r++) o[r] = pack_chan<NTAB>(X, h, r / 2);
        double old[2][4];
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
            const int dd = min(d0 + hh, nd - 1);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double* dst = o[r] <= 0 ? pot : (o[r] <= 2 ? grad : hess);
                const int str = o[r] <= 0 ? 1 : (o[r] <= 2 ? 2 : 3);
                const int cmp = o[r] <= 0 ? 0 : (o[r] <= 2 ? o[r] - 1 : o[r] - 3);
                const size_t pb = (size_t)b * nn + cl + (size_t)n * min(4 * (r % 2) + q, n - 1);
                old[hh][r] = (o[r] >= 0 && !fresh) ? dst[(pb * str + cmp) * nd + dd] : 0.0;
            }
        }
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
            const int dd = d0 + hh;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double* dst = o[r] <= 0 ? pot : (o[r] <= 2 ? grad : hess);
                const int str = o[r] <= 0 ? 1 : (o[r] <= 2 ? 2 : 3);
                const int cmp = o[r] <= 0 ? 0 : (o[r] <= 2 ? o[r] - 1 : o[r] - 3);
                const int k2 = 4 * (r % 2) + q;
                if (o[r] >= 0 && dd < nd && k2 < n && cc < n) {
                    const size_t pb = (size_t)b * nn + cc + (size_t)n * k2;
                    dst[(pb * str + cmp) * nd + dd] = old[hh][r] + oscale * out[hh][r];
                }
            }
        }
    }
}

// Adjoint, one wave per (source, density) over the (target, offsets) pairs
// that read it:
//   W_v   = sum_{o: x(o) = v} G_o Ty_{y(o)}   (K = target y index)
//   fbar += sum_v Tx_v^T W_v                (K = target x index)
// with element s of W_v's accumulator the B operand of K-step s.
// Split-FP32 counterpart of k_near_pack (n <= 8, two or more densities):
// the same tiles, x contraction FF_X[(h, j2)][(g, k1)] on the FP32 matrix
// cores with hi/lo operands.  The FP32 accumulator holds row 4q + r of FF in
// element r of lane (q, c) -- density half q / 2, source y index
// 4 (q % 2) + r -- so the y contraction of density hh takes K-step r with K
// index 4q + r and an A operand that is zero on the lanes of the other
// density half (two products per density instead of the FP64 kernel's split
// of the K-steps between the densities).  The output accumulator's element r
// of lane (q, c) is row 4q + r: y-table half q / 2, target y index
// 4 (q % 2) + r.
template <int NKS, int NTAB>
__global__ __launch_bounds__(64 * kNearWaves) void k_near_pack_split(
    int n, int nd, int nitems, const int* __restrict__ tgt, const int* __restrict__ tlev,
    const int* __restrict__ ptr, const NearPair* __restrict__ pairs, const double* __restrict__ tnearT,
    const double* __restrict__ fvals, double* __restrict__ pot, double* __restrict__ grad,
    double* __restrict__ hess, double oscale, const int* __restrict__ tfresh, int accum) {
    static_assert(NKS <= 2, "two densities per tile need n <= 8");
    constexpr int NX = NTAB == 3 ? 2 : 1;  // tiles, one wave each
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kNearWaves + wave);
    if (item >= nitems) return;  // no block-wide barriers in this kernel
    const int ndp = (nd + 1) / 2;
    const int X = item % NX, it = item / NX / ndp, d0 = 2 * (item / NX % ndp);
    const int q = lane >> 4, c = lane & 15, h = c >> 3, cc = c & 7;
    const int b = tgt[it];
    const bool fresh = accum == 0 && tfresh[it] != 0;  // as in k_near
    const int p0 = ptr[it], p1 = ptr[it + 1];
    const double* T = tnearT + (size_t)tlev[it] * 13 * 3 * 256;
    const int nn = n * n;
    const int cl = min(cc, n - 1);
    const int dh = min(d0 + h, nd - 1);
    const int tx = h ? pack_tx<NTAB>(X, 1) : pack_tx<NTAB>(X, 0);
    const int ty = h ? pack_ty<NTAB>(X, 1) : pack_ty<NTAB>(X, 0);
    const int qh = q >> 1;  // density half of the FF rows this lane holds
    v4f out[2] = {v4f{0, 0, 0, 0}, v4f{0, 0, 0, 0}};
    for (int p = p0; p < p1; p += 2) {
        const bool two = p + 1 < p1;
        const NearPair pr[2] = {pairs[p], pairs[two ? p + 1 : p]};
        v4f FF[2];
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const double* fs = fvals + (size_t)pr[s2].src * nn * nd + dh;
            const double* Tx = T + (size_t)(pr[s2].ix + 6) * 3 * 256;
            const double wgt = (s2 == 0 || two) ? 1.0 : 0.0;
            FF[s2] = v4f{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < NKS; s++) {
                const int j1 = 4 * s + q;
                const Split a = split(wgt * fs[(size_t)(min(j1, n - 1) + n * cl) * nd]);
                FF[s2] = mfma3(a, split(Tx[tx * 256 + j1 * 16 + cc]), FF[s2]);
            }
        }
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const double* Ty = T + (size_t)(pr[s2].iy + 6) * 3 * 256;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j2 = 4 * (q & 1) + r;
                const Split ya = split(Ty[ty * 256 + j2 * 16 + cc]);  // A[(v, k2 = cc)][j2]
                const Split z = Split{0.0f, 0.0f};
#pragma unroll
                for (int hh = 0; hh < 2; hh++) out[hh] = mfma2(qh == hh ? ya : z, FF[s2][r], out[hh]);
            }
        }
    }
    if (p1 > p0) {
        const int o = pack_chan<NTAB>(X, h, qh);
        double* dst = o <= 0 ? pot : (o <= 2 ? grad : hess);
        const int str = o <= 0 ? 1 : (o <= 2 ? 2 : 3);
        const int cmp = o <= 0 ? 0 : (o <= 2 ? o - 1 : o - 3);
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
            const int dd = d0 + hh;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k2 = 4 * (q & 1) + r;
                if (o >= 0 && dd < nd && k2 < n && cc < n) {
                    double& v = dst[(((size_t)b * nn + cc + (size_t)n * k2) * str + cmp) * nd + dd];
                    v = (fresh ? 0.0 : v) + oscale * (double)out[hh][r];
                }
            }
        }
    }
}

template <int NKS, int NTAB, int NPP>
__global__ __launch_bounds__(64 * kNearWaves) void k_near_adj_mfma(
    int n, int nd, int nitems, const int* __restrict__ src, const int* __restrict__ ptr,
    const NearPairT* __restrict__ pairs, const double* __restrict__ tnear, const double* __restrict__ gpot,
    const double* __restrict__ ggrad, const double* __restrict__ ghess, double* __restrict__ fbar) {
    constexpr int NOUT = near_chan(NTAB);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    // wave-uniform: schedule entries are then read with scalar loads
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kNearWaves + wave);
    if (item >= nitems) return;  // no block-wide barriers in this kernel
    const int is = item / nd, d = item % nd;
    const int q = lane >> 4, c = lane & 15;
    const int sb = src[is];
    const int p0 = ptr[is], p1 = ptr[is + 1];
    const int nn = n * n;
    const int cl = min(c, n - 1);
    // NPP pairs per pass so that their loads overlap (a pass past the end of
    // the list repeats its last pair with weight zero), one accumulator each
    v4d acc2[NPP];
#pragma unroll
    for (int h = 0; h < NPP; h++) acc2[h] = v4d{0, 0, 0, 0};
    for (int p = p0; p < p1; p += NPP) {
        NearPairT prs[NPP];
#pragma unroll
        for (int h = 0; h < NPP; h++) prs[h] = pairs[min(p + h, p1 - 1)];
        v4d W[NPP][NTAB];
#pragma unroll
        for (int h = 0; h < NPP; h++) {
            const NearPairT pr = prs[h];
            const double wgt = p + h < p1 ? 1.0 : 0.0;
            const double* T = tnear + (size_t)pr.lev * 13 * 3 * 256;
            const double* Ty = T + (size_t)(pr.iy + 6) * 3 * 256;
            const size_t tb = (size_t)pr.tgt * nn;
#pragma unroll
            for (int v = 0; v < NTAB; v++) W[h][v] = v4d{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < NKS; s++) {
                const int k2 = 4 * s + q;
                // A = G_o: A[k1 = c][k2]; B = Ty_u: B[k2][j2 = c] = Ty_u[k2][c]
                const size_t pb = tb + cl + (size_t)n * min(k2, n - 1);
                double ty[NTAB];
#pragma unroll
                for (int u = 0; u < NTAB; u++) ty[u] = Ty[u * 256 + k2 * 16 + c];
#pragma unroll
                for (int o = 0; o < NOUT; o++) {
                    double g;
                    switch (o) {
                        case 0: g = gpot[pb * nd + d]; break;
                        case 1: g = ggrad[(pb * 2 + 0) * nd + d]; break;
                        case 2: g = ggrad[(pb * 2 + 1) * nd + d]; break;
                        case 3: g = ghess[(pb * 3 + 0) * nd + d]; break;
                        case 4: g = ghess[(pb * 3 + 1) * nd + d]; break;
                        default: g = ghess[(pb * 3 + 2) * nd + d]; break;
                    }
                    W[h][near_x(o)] = mfma(wgt * g, ty[near_y(o)], W[h][near_x(o)]);
                }
            }
        }
        // W_v: element s of lane (q, c) = W_v[k1 = 4s + q][j2 = c]
#pragma unroll
        for (int h = 0; h < NPP; h++) {
            const NearPairT pr = prs[h];
            const double* Tx = tnear + (size_t)pr.lev * 13 * 3 * 256 + (size_t)(pr.ix + 6) * 3 * 256;
#pragma unroll
            for (int s = 0; s < NKS; s++) {
                const int k1 = 4 * s + q;
#pragma unroll
                for (int v = 0; v < NTAB; v++)
                    acc2[h] = mfma(Tx[v * 256 + k1 * 16 + c], W[h][v][s], acc2[h]);  // A[j1 = c][k1]
            }
        }
    }
    v4d acc = acc2[0];
#pragma unroll
    for (int h = 1; h < NPP; h++) acc += acc2[h];
    // element r of lane (q, c) = fbar[j1 = 4r + q][j2 = c]
    if (c < n && p1 > p0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int j1 = 4 * r + q;
            if (j1 < n) fbar[((size_t)sb * nn + j1 + (size_t)n * c) * nd + d] += acc[r];
        }
    }
}

// Zero-padded copies of the masked near-field tables, [level][13][3][16][16]
// as [target][source] (tnear) and [source][target] (tnearT).
__global__ void k_pad_near(int n, int nlev, const double* __restrict__ tloc, double* __restrict__ tnear,
                           double* __restrict__ tnearT) {
    const size_t tot = (size_t)nlev * 13 * 3 * 256;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const size_t blk = e / 256;
        const int r = (int)(e % 256), k = r / 16, j = r % 16;
        const double v = (k < n && j < n) ? tloc[blk * n * n + k * n + j] : 0.0;
        tnear[e] = v;
        tnearT[blk * 256 + j * 16 + k] = v;
    }
}

template <int NKS, int NTAB, int NSPLIT>
int launch_near_mfma3(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    int nitems = p.nnear_t * p.nd * NSPLIT;
    int blocks = (nitems + kNearWaves - 1) / kNearWaves;
    k_near_mfma<NKS, NTAB, NSPLIT><<<blocks, 64 * kNearWaves, 0, p.stream>>>(
        p.n, p.nd, nitems, p.d_near_tgt, p.d_near_lev, p.d_near_ptr, p.d_near, p.d_tnearT, fvals, pot, grad, hess,
        p.out_scale, p.d_near_fresh, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// One density gives one wave per target, too few to fill the chip: its
// tables are then split over two waves (x table 0 | the others; c3 tree,
// n = 8, pot + grad 16.3 -> 13.3 us; n = 16, three tables 22.6 -> 17.7 us).
// With more densities the split's second read of the leaf values costs more
// than it gains (n = 16, two densities: 24.9 -> 28.7 us).
template <int NKS, int NTAB>
int launch_near_mfma2(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    if constexpr (NTAB >= 2) {
        if (p.nd == 1) return launch_near_mfma3<NKS, NTAB, 2>(p, fvals, pot, grad, hess);
    }
    return launch_near_mfma3<NKS, NTAB, 1>(p, fvals, pot, grad, hess);
}

template <int NKS>
int launch_near_mfma1(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    if (p.ifpgh == 1) return launch_near_mfma2<NKS, 1>(p, fvals, pot, grad, hess);
    if (p.ifpgh == 2) return launch_near_mfma2<NKS, 2>(p, fvals, pot, grad, hess);
    return launch_near_mfma2<NKS, 3>(p, fvals, pot, grad, hess);
}

template <int NKS, int NTAB>
int launch_near_pack2(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    int nitems = p.nnear_t * ((p.nd + 1) / 2) * (NTAB == 3 ? 2 : 1);
    int blocks = (nitems + kNearWaves - 1) / kNearWaves;
    k_near_pack<NKS, NTAB><<<blocks, 64 * kNearWaves, 0, p.stream>>>(p.n, p.nd, nitems, p.d_near_tgt, p.d_near_lev,
                                                                   p.d_near_ptr, p.d_near, p.d_tnearT, fvals, pot,
                                                                   grad, hess, p.out_scale, p.d_near_fresh, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NTAB>
int launch_near_pack_split(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    int nitems = p.nnear_t * ((p.nd + 1) / 2) * (NTAB == 3 ? 2 : 1);
    int blocks = (nitems + kNearWaves - 1) / kNearWaves;
    auto kern = p.n <= 4 ? k_near_pack_split<1, NTAB> : k_near_pack_split<2, NTAB>;
    kern<<<blocks, 64 * kNearWaves, 0, p.stream>>>(p.n, p.nd, nitems, p.d_near_tgt, p.d_near_lev, p.d_near_ptr,
                                                   p.d_near, p.d_tnearT, fvals, pot, grad, hess, p.out_scale,
                                                   p.d_near_fresh, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NKS, int NTAB>
int launch_near_split2(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    int nitems = p.nnear_t * p.nd;
    int blocks = (nitems + kNearWaves - 1) / kNearWaves;
    k_near_split<NKS, NTAB><<<blocks, 64 * kNearWaves, 0, p.stream>>>(p.n, p.nd, nitems, p.d_near_tgt, p.d_near_lev,
                                                                    p.d_near_ptr, p.d_near, p.d_tnearT, fvals, pot,
                                                                    grad, hess, p.out_scale, p.d_near_fresh, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NKS>
int launch_near_split1(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    if (p.ifpgh == 1) return launch_near_split2<NKS, 1>(p, fvals, pot, grad, hess);
    if (p.ifpgh == 2) return launch_near_split2<NKS, 2>(p, fvals, pot, grad, hess);
    return launch_near_split2<NKS, 3>(p, fvals, pot, grad, hess);
}

template <int NKS>
int launch_near_pack1(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    if (p.ifpgh == 1) return launch_near_pack2<NKS, 1>(p, fvals, pot, grad, hess);
    if (p.ifpgh == 2) return launch_near_pack2<NKS, 2>(p, fvals, pot, grad, hess);
    return launch_near_pack2<NKS, 3>(p, fvals, pot, grad, hess);
}

// One density gives one wave per source, too few to hide the loads of a
// pair behind the products of another: two pairs per pass then (c3 tree with
// one density 33.2 -> 28.8 us); with three densities one (53.0 -> 54.7 us
// with two).
template <int NKS, int NTAB>
int launch_near_adj_mfma2(Plan& p, const double* gpot, const double* ggrad, const double* ghess, double* fbar) {
    int nitems = p.nnearT_s * p.nd;
    int blocks = (nitems + kNearWaves - 1) / kNearWaves;
    if (p.nd == 1) {
        k_near_adj_mfma<NKS, NTAB, 2><<<blocks, 64 * kNearWaves, 0, p.stream>>>(
            p.n, p.nd, nitems, p.d_nearT_src, p.d_nearT_ptr, p.d_nearT, p.d_tnear, gpot, ggrad, ghess, fbar);
        BFGT_HIP_CHECK(hipGetLastError());
        return BFGT_OK;
    }
    k_near_adj_mfma<NKS, NTAB, 1><<<blocks, 64 * kNearWaves, 0, p.stream>>>(
        p.n, p.nd, nitems, p.d_nearT_src, p.d_nearT_ptr, p.d_nearT, p.d_tnear, gpot, ggrad, ghess, fbar);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NKS>
int launch_near_adj_mfma1(Plan& p, const double* gpot, const double* ggrad, const double* ghess, double* fbar) {
    if (p.ifpgh == 1) return launch_near_adj_mfma2<NKS, 1>(p, gpot, ggrad, ghess, fbar);
    if (p.ifpgh == 2) return launch_near_adj_mfma2<NKS, 2>(p, gpot, ggrad, ghess, fbar);
    return launch_near_adj_mfma2<NKS, 3>(p, gpot, ggrad, ghess, fbar);
}

}  // namespace

bool near_use_mfma(const Plan& p) { return p.n <= 16; }

int build_near_padded(Plan& p) {
    if (!near_use_mfma(p)) return BFGT_OK;
    int nlev = p.nlevels + 1;
    size_t tot = (size_t)nlev * 13 * 3 * 256;
    int blocks = (int)((tot + 255) / 256 < 4096 ? (tot + 255) / 256 : 4096);
    k_pad_near<<<blocks, 256, 0, p.stream>>>(p.n, nlev, p.d_tloc, p.d_tnear, p.d_tnearT);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

size_t near_smem_bytes(int n, int ntab) { return sizeof(double) * (size_t)n * n * (1 + ntab); }

int launch_near(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    if (p.nnear_t == 0) return BFGT_OK;
    if (near_use_mfma(p) && p.precision != BFGT_PREC_FP64) {  // split-FP32 for both reduced precisions
        if (p.n <= 8 && p.nd >= 2) {  // two densities per tile, as the FP64 pass
            if (p.ifpgh == 1) return launch_near_pack_split<1>(p, fvals, pot, grad, hess);
            if (p.ifpgh == 2) return launch_near_pack_split<2>(p, fvals, pot, grad, hess);
            return launch_near_pack_split<3>(p, fvals, pot, grad, hess);
        }
        switch ((p.n + 3) / 4) {
            case 1: return launch_near_split1<1>(p, fvals, pot, grad, hess);
            case 2: return launch_near_split1<2>(p, fvals, pot, grad, hess);
            case 3: return launch_near_split1<3>(p, fvals, pot, grad, hess);
            default: return launch_near_split1<4>(p, fvals, pot, grad, hess);
        }
    }
    if (near_use_mfma(p)) {
        if (p.n <= 4 && p.nd >= 2) return launch_near_pack1<1>(p, fvals, pot, grad, hess);
        if (p.n <= 8 && p.nd >= 2) return launch_near_pack1<2>(p, fvals, pot, grad, hess);
        switch ((p.n + 3) / 4) {
            case 1: return launch_near_mfma1<1>(p, fvals, pot, grad, hess);
            case 2: return launch_near_mfma1<2>(p, fvals, pot, grad, hess);
            case 3: return launch_near_mfma1<3>(p, fvals, pot, grad, hess);
            default: return launch_near_mfma1<4>(p, fvals, pot, grad, hess);
        }
    }
    int ntab = p.ifpgh;
    size_t sh = near_smem_bytes(p.n, ntab);
    k_near<<<p.nnear_t, kNearThreads, sh, p.stream>>>(p.n, p.nd, ntab, p.d_near_tgt, p.d_near_lev, p.d_near_ptr,
                                                      p.d_near, p.d_tloc, fvals, pot, grad, hess, p.out_scale, 6,
                                                      p.d_near_fresh, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// Near field of an explicit list of targets (boxes refined below the
// tree's leaves): ntgt targets tgt[i] at levels tlev[i], their (source,
// offsets) pairs in CSR ptr/pairs, tables with offsets -half..half.
int launch_near_list(Plan& p, int ntgt, const int* tgt, const int* tlev, const int* ptr, const NearPair* pairs,
                     const double* tloc, int half, const double* fvals, double* pot, double* grad, double* hess) {
    if (ntgt == 0) return BFGT_OK;
    int ntab = p.ifpgh;
    size_t sh = near_smem_bytes(p.n, ntab);
    k_near<<<ntgt, kNearThreads, sh, p.stream>>>(p.n, p.nd, ntab, tgt, tlev, ptr, pairs, tloc, fvals, pot, grad,
                                                 hess, 1.0, half, nullptr, 1);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_near_adj(Plan& p, const double* gpot, const double* ggrad, const double* ghess, double* fbar) {
    if (p.nnearT_s == 0) return BFGT_OK;
    if (near_use_mfma(p)) {
        switch ((p.n + 3) / 4) {
            case 1: return launch_near_adj_mfma1<1>(p, gpot, ggrad, ghess, fbar);
            case 2: return launch_near_adj_mfma1<2>(p, gpot, ggrad, ghess, fbar);
            case 3: return launch_near_adj_mfma1<3>(p, gpot, ggrad, ghess, fbar);
            default: return launch_near_adj_mfma1<4>(p, gpot, ggrad, ghess, fbar);
        }
    }
    int ntab = p.ifpgh;
    int nout = ntab == 1 ? 1 : (ntab == 2 ? 3 : 6);
    size_t sh = sizeof(double) * (size_t)p.n * p.n * (nout + ntab);
    if (sh > 65536)
        BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)k_near_adj, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    k_near_adj<<<p.nnearT_s, kNearThreads, sh, p.stream>>>(p.n, p.nd, ntab, p.d_nearT_src, p.d_nearT_ptr, p.d_nearT,
                                                           p.d_tloc, gpot, ggrad, ghess, fbar);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

}  // namespace bfgt
