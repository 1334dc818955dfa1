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
                                                       double* __restrict__ grad, double* __restrict__ hess) {
    extern __shared__ double smem[];
    const int nn = n * n;
    double* sf = smem;       // nn
    double* sff = sf + nn;   // ntab * nn : [t][k1 + n*j2]
    const int b = tgt[blockIdx.x];
    const int lev = tlev[blockIdx.x];
    const int p0 = ptr[blockIdx.x], p1 = ptr[blockIdx.x + 1];
    const int tid = threadIdx.x;
    const double* T = tloc + (size_t)lev * 13 * 3 * nn;
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
            const double* Tx = T + (size_t)(pr.ix + 6) * 3 * nn;
            const double* Ty = T + (size_t)(pr.iy + 6) * 3 * nn;
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
                pot[pb * nd + d] += acc[r][0];
                if (ntab >= 2) {
                    grad[(pb * 2 + 0) * nd + d] += acc[r][1];
                    grad[(pb * 2 + 1) * nd + d] += acc[r][2];
                }
                if (ntab >= 3) {
                    hess[(pb * 3 + 0) * nd + d] += acc[r][3];
                    hess[(pb * 3 + 1) * nd + d] += acc[r][4];
                    hess[(pb * 3 + 2) * nd + d] += acc[r][5];
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

}  // namespace

size_t near_smem_bytes(int n, int ntab) { return sizeof(double) * (size_t)n * n * (1 + ntab); }

int launch_near(Plan& p, const double* fvals, double* pot, double* grad, double* hess) {
    if (p.nnear_t == 0) return BFGT_OK;
    int ntab = p.ifpgh;
    size_t sh = near_smem_bytes(p.n, ntab);
    k_near<<<p.nnear_t, kNearThreads, sh, p.stream>>>(p.n, p.nd, ntab, p.d_near_tgt, p.d_near_lev, p.d_near_ptr,
                                                      p.d_near, p.d_tloc, fvals, pot, grad, hess);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_near_adj(Plan& p, const double* gpot, const double* ggrad, const double* ghess, double* fbar) {
    if (p.nnearT_s == 0) return BFGT_OK;
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
