// Plane-wave stages of the box FGT on the GPU (FP64).
//
// The reference forms a PW expansion per leaf (bfgt/boxfgt.f:728-744), merges
// children into parents level by level (:764-783), translates at the cutoff
// level (:805-822), splits parents into children level by level (:845-863)
// and evaluates on leaves (:884-900).  Every merge/split factor is a diagonal
// phase, so here the leaf expansion is shifted straight to its ancestor at
// the cutoff level and the ancestor's local expansion straight back to the
// leaf; the per-level expansions of interior boxes are never stored.  The
// phase is separable (x frequency times y frequency) and is folded into the
// 1-D contractions.  Only leaf values, the cutoff-level expansions and the
// outputs touch HBM.
//
// Expansion layout: [slot][density][k2][k1] complex, k1 < npw the x
// frequency, k2 < H = (npw+1)/2 the retained half of the y frequencies
// (pwexp(npw, H, nd), bfgt/bfgt_pwrouts.f:98-147).
#include <hip/hip_runtime.h>

#include "bfgt_internal.h"

namespace bfgt {

namespace {

__device__ inline cplx cmul(cplx a, cplx b) { return cplx{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline cplx cfma(cplx a, cplx b, cplx c) {
    return cplx{fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
__device__ inline cplx cexpi(double t) {
    double s, c;
    sincos(t, &s, &c);
    return cplx{c, s};
}

constexpr int kUpThreads = 256;
constexpr int kUpMaxOut = 32 * kUpThreads;  // largest nexp the upward kernel accepts

// Leaf values -> PW expansions about the cutoff-level ancestor, summed over
// the leaves of one chunk.  One block per chunk; each thread accumulates
// MAXE expansion coefficients in registers.
template <int kUpMaxE>
__global__ __launch_bounds__(kUpThreads) void k_up(int n, int npw, int nd, const double* __restrict__ fvals,
                                                   const PwLeaf* __restrict__ leaves,
                                                   const UpChunk* __restrict__ chunks,
                                                   const cplx* __restrict__ tpoly, const double* __restrict__ ts,
                                                   cplx* __restrict__ partial) {
    extern __shared__ double smem[];
    const int H = (npw + 1) / 2;
    const int nexp = npw * H;
    const int npbox = n * n;
    double* sf = smem;                           // n*n
    cplx* sT = reinterpret_cast<cplx*>(sf + npbox);  // npw*n : sT[k*n + m]
    cplx* sff = sT + npw * n;                    // n*H : sff[m1 + n*k2]
    cplx* spx = sff + n * H;                     // npw
    cplx* spy = spx + npw;                       // H
    const UpChunk ch = chunks[blockIdx.x];
    const int tid = threadIdx.x;
    for (int d = 0; d < nd; d++) {
        cplx acc[kUpMaxE];
#pragma unroll
        for (int r = 0; r < kUpMaxE; r++) acc[r] = cplx{0.0, 0.0};
        int cur_level = -1;
        for (int li = 0; li < ch.count; li++) {
            const PwLeaf lf = leaves[ch.first + li];
            __syncthreads();
            if (lf.level != cur_level) {
                const cplx* src = tpoly + (size_t)lf.level * npw * n;
                for (int e = tid; e < npw * n; e += kUpThreads) sT[e] = src[e];
                cur_level = lf.level;
            }
            const double* fb = fvals + (size_t)lf.box * npbox * nd;
            for (int e = tid; e < npbox; e += kUpThreads) sf[e] = fb[(size_t)e * nd + d];
            for (int k = tid; k < npw; k += kUpThreads) spx[k] = cexpi(-ts[k] * lf.dx);
            for (int k = tid; k < H; k += kUpThreads) spy[k] = cexpi(-ts[k] * lf.dy);
            __syncthreads();
            // y-transform to the half spectrum, y phase folded in
            for (int e = tid; e < n * H; e += kUpThreads) {
                int m1 = e % n, k2 = e / n;
                double sr = 0.0, si = 0.0;
                const cplx* trow = sT + k2 * n;
                for (int m2 = 0; m2 < n; m2++) {
                    double f = sf[m1 + n * m2];
                    sr = fma(trow[m2].x, f, sr);
                    si = fma(trow[m2].y, f, si);
                }
                sff[e] = cmul(cplx{sr, si}, spy[k2]);
            }
            __syncthreads();
            // x-transform to the full spectrum, x phase folded in
#pragma unroll
            for (int r = 0; r < kUpMaxE; r++) {
                int e = tid + r * kUpThreads;
                if (e < nexp) {
                    int k1 = e % npw, k2 = e / npw;
                    const cplx* trow = sT + k1 * n;
                    const cplx* fcol = sff + n * k2;
                    cplx s{0.0, 0.0};
                    for (int m1 = 0; m1 < n; m1++) s = cfma(trow[m1], fcol[m1], s);
                    acc[r] = cfma(s, spx[k1], acc[r]);
                }
            }
        }
        cplx* out = partial + ((size_t)ch.slot * nd + d) * nexp;
#pragma unroll
        for (int r = 0; r < kUpMaxE; r++) {
            int e = tid + r * kUpThreads;
            if (e < nexp) out[e] = acc[r];
        }
    }
}

// mp of each cutoff-level box = sum of its chunks' partial sums (in chunk order).
__global__ void k_reduce(int nd, int nexp, const int* __restrict__ anc_chunk, const cplx* __restrict__ partial,
                         cplx* __restrict__ mp) {
    const int a = blockIdx.x;
    const int len = nd * nexp;
    const int c0 = anc_chunk[a], c1 = anc_chunk[a + 1];
    for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < len; e += gridDim.y * blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int c = c0; c < c1; c++) {
            cplx v = partial[(size_t)c * len + e];
            sr += v.x;
            si += v.y;
        }
        mp[(size_t)a * len + e] = cplx{sr, si};
    }
}

// loc = mp(self) + sum over the PW list of shift(offset) * mp(source)
// (bfgt/boxfgt.f:805-822).
__global__ void k_m2l(int nd, int nexp, const int* __restrict__ m2l_ptr, const int2* __restrict__ m2l,
                      const cplx* __restrict__ wshift, const cplx* __restrict__ mp, cplx* __restrict__ loc) {
    const int a = blockIdx.x;
    const int len = nd * nexp;
    const int j0 = m2l_ptr[a], j1 = m2l_ptr[a + 1];
    for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < len; e += gridDim.y * blockDim.x) {
        int q = e % nexp;
        cplx s = mp[(size_t)a * len + e];
        for (int j = j0; j < j1; j++) {
            int2 sj = m2l[j];
            s = cfma(wshift[(size_t)sj.y * nexp + q], mp[(size_t)sj.x * len + e], s);
        }
        loc[(size_t)a * len + e] = s;
    }
}

constexpr int kDownThreads = 256;
constexpr int kDownChunk = 8;  // y-frequency columns staged in LDS at a time

// Local expansion of the cutoff-level ancestor -> pot/grad/hess on one leaf
// (split + evaluation, bfgt/boxfgt.f:845-900, bfgt/bfgt_pwrouts.f:549-784).
// The expansion is streamed through LDS kDownChunk y-frequency columns at a
// time; each thread keeps the running sums of its NQ output points.
template <int NQ>
__global__ __launch_bounds__(kDownThreads) void k_down(int n, int npw, int nd, int ntab,
                                                       const PwLeaf* __restrict__ leaves,
                                                       const cplx* __restrict__ tpw, const double* __restrict__ ts,
                                                       const cplx* __restrict__ loc, double* __restrict__ pot,
                                                       double* __restrict__ grad, double* __restrict__ hess) {
    extern __shared__ double smem[];
    const int H = (npw + 1) / 2;
    const int npw2 = npw / 2;
    const int nexp = npw * H;
    const int npbox = n * n;
    cplx* sE = reinterpret_cast<cplx*>(smem);  // ntab*npw*n : [t][k][i]
    cplx* sL = sE + ntab * npw * n;            // npw*kDownChunk : [mm][m1]
    cplx* sff = sL + npw * kDownChunk;         // ntab*kDownChunk*n : [t][mm][k1]
    cplx* spx = sff + ntab * kDownChunk * n;   // npw
    cplx* spy = spx + npw;                     // H
    const PwLeaf lf = leaves[blockIdx.x];
    const int tid = threadIdx.x;
    const cplx* E = tpw + (size_t)lf.level * 3 * npw * n;
    for (int e = tid; e < ntab * npw * n; e += kDownThreads) sE[e] = E[e];
    for (int k = tid; k < npw; k += kDownThreads) spx[k] = cexpi(ts[k] * lf.dx);
    for (int k = tid; k < H; k += kDownThreads) spy[k] = cexpi(ts[k] * lf.dy);
    for (int d = 0; d < nd; d++) {
        double acc[NQ][6];
#pragma unroll
        for (int r = 0; r < NQ; r++)
#pragma unroll
            for (int c = 0; c < 6; c++) acc[r][c] = 0.0;
        const cplx* L = loc + ((size_t)lf.anc * nd + d) * nexp;
        for (int c0 = 0; c0 < H; c0 += kDownChunk) {
            const int mc = min(kDownChunk, H - c0);
            __syncthreads();
            for (int e = tid; e < npw * mc; e += kDownThreads) {
                int m1 = e % npw, mm = e / npw;
                sL[mm * npw + m1] = cmul(L[(size_t)(c0 + mm) * npw + m1], spx[m1]);
            }
            __syncthreads();
            // x contraction of the staged columns, y phase folded in
            for (int e = tid; e < ntab * mc * n; e += kDownThreads) {
                int t = e / (mc * n), r = e % (mc * n);
                int mm = r / n, k1 = r % n;
                const cplx* Et = sE + t * npw * n;
                const cplx* Lc = sL + mm * npw;
                cplx s{0.0, 0.0};
                for (int m1 = 0; m1 < npw; m1++) s = cfma(Et[m1 * n + k1], Lc[m1], s);
                sff[(t * kDownChunk + mm) * n + k1] = cmul(s, spy[c0 + mm]);
            }
            __syncthreads();
            // y contraction into the per-point sums (real parts; the middle
            // frequency of an odd expansion counts half)
#pragma unroll
            for (int r = 0; r < NQ; r++) {
                int q = tid + r * kDownThreads;
                if (q < npbox) {
                    int k1 = q % n, k2 = q / n;
                    for (int mm = 0; mm < mc; mm++) {
                        int m2 = c0 + mm;
                        double w = (m2 < npw2) ? 1.0 : 0.5;
                        cplx e0 = sE[(0 * npw + m2) * n + k2];
                        cplx f0 = sff[(0 * kDownChunk + mm) * n + k1];
                        acc[r][0] += w * (e0.x * f0.x - e0.y * f0.y);
                        if (ntab >= 2) {
                            cplx e1 = sE[(1 * npw + m2) * n + k2];
                            cplx f1 = sff[(1 * kDownChunk + mm) * n + k1];
                            acc[r][1] += w * (e0.x * f1.x - e0.y * f1.y);
                            acc[r][2] += w * (e1.x * f0.x - e1.y * f0.y);
                            if (ntab >= 3) {
                                cplx e2 = sE[(2 * npw + m2) * n + k2];
                                cplx f2 = sff[(2 * kDownChunk + mm) * n + k1];
                                acc[r][3] += w * (e0.x * f2.x - e0.y * f2.y);
                                acc[r][4] += w * (e1.x * f1.x - e1.y * f1.y);
                                acc[r][5] += w * (e2.x * f0.x - e2.y * f0.y);
                            }
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < NQ; r++) {
            int q = tid + r * kDownThreads;
            if (q < npbox) {
                size_t pb = (size_t)lf.box * npbox + q;
                pot[pb * nd + d] += 2.0 * acc[r][0];
                if (ntab >= 2) {
                    grad[(pb * 2 + 0) * nd + d] += 2.0 * acc[r][1];
                    grad[(pb * 2 + 1) * nd + d] += 2.0 * acc[r][2];
                }
                if (ntab >= 3) {
                    hess[(pb * 3 + 0) * nd + d] += 2.0 * acc[r][3];
                    hess[(pb * 3 + 1) * nd + d] += 2.0 * acc[r][4];
                    hess[(pb * 3 + 2) * nd + d] += 2.0 * acc[r][5];
                }
            }
        }
    }
}

}  // namespace

size_t up_smem_bytes(int n, int npw) {
    int H = (npw + 1) / 2;
    return sizeof(double) * n * n + sizeof(cplx) * ((size_t)npw * n + (size_t)n * H + npw + H);
}

size_t down_smem_bytes(int n, int npw, int ntab) {
    int H = (npw + 1) / 2;
    return sizeof(cplx) * ((size_t)ntab * npw * n + (size_t)npw * kDownChunk + (size_t)ntab * kDownChunk * n + npw + H);
}

int up_max_out() { return kUpMaxOut; }

template <int E>
static int launch_up_t(Plan& p, const double* fvals, size_t sh) {
    if (sh > 65536)
        BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)k_up<E>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    k_up<E><<<p.nchunks, kUpThreads, sh, p.stream>>>(p.n, p.npw, p.nd, fvals, p.d_pwleaf, p.d_chunks, p.d_tpoly,
                                                     p.d_ts, p.d_partial);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_upward(Plan& p, const double* fvals) {
    if (p.nchunks == 0) return BFGT_OK;
    size_t sh = up_smem_bytes(p.n, p.npw);
    int e = (p.nexp + kUpThreads - 1) / kUpThreads;
    if (e <= 2) return launch_up_t<2>(p, fvals, sh);
    if (e <= 4) return launch_up_t<4>(p, fvals, sh);
    if (e <= 8) return launch_up_t<8>(p, fvals, sh);
    if (e <= 16) return launch_up_t<16>(p, fvals, sh);
    if (e <= 32) return launch_up_t<32>(p, fvals, sh);
    return set_error(BFGT_ERR_RESOURCE, "plane-wave expansion too long for the leaf kernels");
}

int launch_reduce(Plan& p) {
    int len = p.nd * p.nexp;
    int gy = (len + 255) / 256;
    k_reduce<<<dim3(p.nanc, gy), 256, 0, p.stream>>>(p.nd, p.nexp, p.d_anc_chunk, p.d_partial, p.d_mp);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_m2l(Plan& p) {
    int len = p.nd * p.nexp;
    int gy = (len + 255) / 256;
    k_m2l<<<dim3(p.nanc, gy), 256, 0, p.stream>>>(p.nd, p.nexp, p.d_m2l_ptr, p.d_m2l, p.d_wshift, p.d_mp, p.d_loc);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NQ>
static int launch_down_t(Plan& p, double* pot, double* grad, double* hess, size_t sh) {
    if (sh > 65536)
        BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)k_down<NQ>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    k_down<NQ><<<p.npwleaf, kDownThreads, sh, p.stream>>>(p.n, p.npw, p.nd, p.ifpgh, p.d_pwleaf, p.d_tpw, p.d_ts,
                                                          p.d_loc, pot, grad, hess);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_downward(Plan& p, double* pot, double* grad, double* hess) {
    if (p.npwleaf == 0) return BFGT_OK;
    size_t sh = down_smem_bytes(p.n, p.npw, p.ifpgh);
    int q = (p.n * p.n + kDownThreads - 1) / kDownThreads;
    if (q <= 1) return launch_down_t<1>(p, pot, grad, hess, sh);
    if (q <= 2) return launch_down_t<2>(p, pot, grad, hess, sh);
    return launch_down_t<4>(p, pot, grad, hess, sh);
}

}  // namespace bfgt
