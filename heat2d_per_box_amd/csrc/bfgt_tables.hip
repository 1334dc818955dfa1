// FP64 table generation on the GPU for the box FGT.
//
// Every 1-D table the transform needs is built here, in double precision,
// from the node rules up:
//   * Gauss-Legendre / Chebyshev node rules and the values->coefficients
//     matrices U (common/legeexps.f:140-220, common/chebexps.f:12-180);
//   * values->plane-wave tables via spherical Bessel functions
//     (bfgt/bfgt_volrouts.f:58-275);
//   * plane-wave->pot/grad/hess tables (bfgt/bfgt_volrouts.f:339-401);
//   * near-field tables of the 13 geometric offsets per target level, the
//     erf-recurrence / 50-point quadrature integrals folded with U, their
//     mirror images and eps-sparsity masks (bfgt/bfgt_volrouts.f:407-660,
//     920-1378);
//   * the nine mp->loc diagonal shifts at the cutoff level
//     (bfgt/bfgt_pwrouts.f:1415-1471).
// Sparsity is applied by zeroing the entries outside the reference's index
// ranges, so a dense contraction over a masked table sums exactly the terms
// the reference's sparse loops sum.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>

#include "bfgt_internal.h"

namespace bfgt {

namespace {

// d_nodes layout (doubles)
struct NodeLayout {
    int n;
    __host__ __device__ int xq() const { return 0; }
    __host__ __device__ int U() const { return n; }
    __host__ __device__ int xl() const { return n + n * n; }
    __host__ __device__ int wl() const { return 2 * n + n * n; }
    __host__ __device__ int Ul() const { return 3 * n + n * n; }
    __host__ __device__ int xn() const { return 3 * n + 2 * n * n; }
    __host__ __device__ int wn() const { return 3 * n + 2 * n * n + kNq; }
    __host__ __device__ int size() const { return 3 * n + 2 * n * n + 2 * kNq; }
};

__device__ void legendre_pols(double x, int n, double* p) {
    p[0] = 1.0;
    if (n == 0) return;
    p[1] = x;
    for (int k = 1; k < n; k++) p[k + 1] = ((2 * k + 1) * x * p[k] - k * p[k - 1]) / (k + 1);
}

__device__ void cheb_pols(double x, int n, double* p) {
    p[0] = 1.0;
    if (n == 0) return;
    p[1] = x;
    for (int k = 1; k < n; k++) p[k + 1] = 2 * x * p[k] - p[k - 1];
}

// i-th largest root of P_n and its Gauss weight (Newton from a cosine guess).
__device__ void gl_root(int n, int i, double* xr, double* wr) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double p0 = 1.0, p1 = z, dp = 1.0;
    for (int it = 0; it < 100; it++) {
        p0 = 1.0;
        p1 = z;
        for (int k = 1; k < n; k++) {
            double p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        double dz = p1 / dp;
        z -= dz;
        if (fabs(dz) < 1e-16 && it > 2) break;
    }
    p0 = 1.0;
    p1 = z;
    for (int k = 1; k < n; k++) {
        double p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    *xr = z;
    *wr = 2.0 / ((1.0 - z * z) * dp * dp);
}

// Node rules: one block of 64 threads.
__global__ void k_node_rules(int n, int ipoly, double* out) {
    NodeLayout L{n};
    int t = threadIdx.x;
    // n-point Gauss-Legendre rule (ascending)
    if (t < (n + 1) / 2) {
        double z, w;
        gl_root(n, t, &z, &w);
        if (n % 2 == 1 && t == (n + 1) / 2 - 1) z = 0.0;
        out[L.xl() + n - 1 - t] = z;
        out[L.xl() + t] = -z;
        out[L.wl() + n - 1 - t] = w;
        out[L.wl() + t] = w;
    }
    // 50-point rule for the near-field quadrature branch
    if (t < kNq / 2) {
        double z, w;
        gl_root(kNq, t, &z, &w);
        out[L.xn() + kNq - 1 - t] = z;
        out[L.xn() + t] = -z;
        out[L.wn() + kNq - 1 - t] = w;
        out[L.wn() + t] = w;
    }
    __syncthreads();
    double pols[kMaxOrder + 2];
    if (t < n) {
        // Legendre U: U(i,j) = P_{i}(x_j) w_j (2i+1)/2, stored [i*n + j]
        double xj = out[L.xl() + t], wj = out[L.wl() + t];
        legendre_pols(xj, n - 1, pols);
        for (int i = 0; i < n; i++) out[L.Ul() + i * n + t] = pols[i] * wj * (2 * i + 1) / 2.0;
        if (ipoly == 0) {
            out[L.xq() + t] = xj;
            for (int i = 0; i < n; i++) out[L.U() + i * n + t] = out[L.Ul() + i * n + t];
        } else {
            // Chebyshev nodes cos((2i-1)pi/2n) ascending, U(0,j)=1/n, U(i,j)=2/n T_i(x_j)
            double xc = cos((2 * (n - t) - 1) * M_PI / (2 * n));
            out[L.xq() + t] = xc;
            cheb_pols(xc, n - 1, pols);
            for (int i = 0; i < n; i++) out[L.U() + i * n + t] = (i == 0 ? 1.0 : 2.0) * pols[i] / n;
        }
    }
}

// Spherical Bessel j_0..j_{n-1}(x), x > 0, by Miller's backward recurrence
// normalised with j_0 or j_1 (whichever is larger); O(n) storage.
__device__ void sph_besselj(int n, double x, double* fj) {
    if (fabs(x) < 1e-200) {
        fj[0] = 1.0;
        for (int i = 1; i < n; i++) fj[i] = 0.0;
        return;
    }
    double xinv = 1.0 / x;
    double fm1 = 0.0, f0 = 1.0;
    int ntop = n + 100000;
    for (int i = (n > 0 ? n - 1 : 0); i < n + 100000; i++) {
        double f1 = (2 * i + 1.0) * xinv * f0 - fm1;
        if (f1 * f1 > 1e40) {
            ntop = i + 1;
            break;
        }
        fm1 = f0;
        f0 = f1;
    }
    // backward: a = f_{i+1}, b = f_i
    double a = 0.0, b = 1.0;  // f_ntop, f_{ntop-1}
    for (int k = 0; k < n; k++) fj[k] = 0.0;
    if (ntop - 1 < n) fj[ntop - 1] = b;
    double f1v = 0, f0v = 0;
    for (int i = ntop - 1; i >= 1; i--) {
        double c = (2 * i + 1.0) * xinv * b - a;  // f_{i-1}
        a = b;
        b = c;
        if (i - 1 < n) fj[i - 1] = b;
        if (b * b > 1e40) {
            a *= 1e-20;
            b *= 1e-20;
            for (int k = i - 1; k < n; k++) fj[k] *= 1e-20;
        }
    }
    f0v = b;  // f_0
    f1v = a;  // f_1
    double j0 = sin(x) * xinv;
    double j1 = j0 * xinv - cos(x) * xinv;
    double sc = (fabs(j1) > fabs(j0)) ? j1 / f1v : j0 / f0v;
    for (int k = 0; k < n; k++) fj[k] *= sc;
}

// values -> PW tables, layout [level][m][k] (node m, PW index k fastest).
// grid (npw - npw/2, nlev_pw), block 64.
__global__ void k_poly2pw(int n, int ipoly, int npw, int lev0, const double* ws, const double* ts, double delta0,
                          const double* boxsize, const double* nodes, cplx* tab) {
    NodeLayout L{n};
    int j = npw / 2 + blockIdx.x;
    int lev = lev0 + blockIdx.y;
    double boxdim = boxsize[lev];
    __shared__ double tre[kMaxOrder], tim[kMaxOrder];
    int t = threadIdx.x;
    if (t == 0) {
        double fj[kMaxOrder];
        double dd = ws[j] * boxdim;
        double lre[kMaxOrder], lim[kMaxOrder];
        if (fabs(ts[j]) < 1e-12) {
            for (int m = 0; m < n; m++) {
                if (ipoly == 0) lre[m] = (m == 0) ? dd : 0.0;
                else lre[m] = (m % 2 == 0) ? dd / (1.0 - (double)m * m) : 0.0;
                lim[m] = 0.0;
            }
            for (int m = 0; m < n; m++) {
                tre[m] = lre[m];
                tim[m] = lim[m];
            }
        } else {
            double fac = boxdim / (2 * sqrt(delta0));
            sph_besselj(n, ts[j] * fac, fj);
            // dd * j_m / i^m
            for (int m = 0; m < n; m++) {
                double v = dd * fj[m];
                switch (m & 3) {
                    case 0: lre[m] = v; lim[m] = 0; break;
                    case 1: lre[m] = 0; lim[m] = -v; break;
                    case 2: lre[m] = -v; lim[m] = 0; break;
                    default: lre[m] = 0; lim[m] = v; break;
                }
            }
            if (ipoly == 0) {
                for (int m = 0; m < n; m++) {
                    tre[m] = lre[m];
                    tim[m] = lim[m];
                }
            } else {
                // Chebyshev T_i as a Legendre series: v(k,i) = sum_q Ul(k,q) T_i(xl_q)
                double cp[kMaxOrder + 2];
                for (int i = 0; i < n; i++) {
                    tre[i] = 0;
                    tim[i] = 0;
                }
                for (int q = 0; q < n; q++) {
                    cheb_pols(nodes[L.xl() + q], n - 1, cp);
                    for (int k = 0; k < n; k++) {
                        double u = nodes[L.Ul() + k * n + q];
                        for (int i = 0; i < n; i++) {
                            double v = u * cp[i];
                            tre[i] += v * lre[k];
                            tim[i] += v * lim[k];
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // fold with U: tab(m, j) = sum_k tmp(k) U(k, m); mirror j -> npw-1-j by conjugation
    if (t < n) {
        double sr = 0, si = 0;
        for (int k = 0; k < n; k++) {
            double u = nodes[L.U() + k * n + t];
            sr += tre[k] * u;
            si += tim[k] * u;
        }
        // layout [level][m][k]: the PW index is the fastest
        size_t base = (size_t)lev * npw * n;
        tab[base + (size_t)t * npw + j] = cplx{sr, si};
        int jm = npw - 1 - j;
        if (jm != j) tab[base + (size_t)t * npw + jm] = cplx{sr, -si};
    }
}

// PW -> pot/grad/hess tables.  grid nlev_pw, block 256.
__global__ void k_pw2pgh(int n, int npw, int lev0, const double* ts, const double* boxsize, double delta0,
                         const double* nodes, cplx* tab) {
    NodeLayout L{n};
    int lev = lev0 + blockIdx.x;
    double boxdim = boxsize[lev];
    double dsq = boxdim / 2 / sqrt(delta0);
    double dfac = dsq * 2 / boxdim;  // imaginary unit factor i*dfac
    int npw2 = npw / 2;
    int nup = npw - npw2;
    size_t base = (size_t)lev * 3 * npw * n;
    for (int e = threadIdx.x; e < nup * n; e += blockDim.x) {
        int j = npw2 + e / n, i = e % n;
        double x = nodes[L.xq() + i] * dsq;
        double s, c;
        sincos(ts[j] * x, &s, &c);
        // e = c + i s; ex = ts*e*(i dfac); exx = ts*ex*(i dfac)
        double exr = -ts[j] * dfac * s, exi = ts[j] * dfac * c;
        double exxr = -ts[j] * dfac * exi, exxi = ts[j] * dfac * exr;
        int jm = npw - 1 - j;
        tab[base + (size_t)(0 * npw + j) * n + i] = cplx{c, s};
        tab[base + (size_t)(1 * npw + j) * n + i] = cplx{exr, exi};
        tab[base + (size_t)(2 * npw + j) * n + i] = cplx{exxr, exxi};
        tab[base + (size_t)(0 * npw + jm) * n + i] = cplx{c, -s};
        tab[base + (size_t)(1 * npw + jm) * n + i] = cplx{exr, -exi};
        tab[base + (size_t)(2 * npw + jm) * n + i] = cplx{exxr, -exxi};
    }
}

// int_{-1}^{1} P_m(x) exp(-(t-x)^2/lambda^2) dx and t-derivatives by the
// five-term recurrence (bfgt/bfgt_volrouts.f:1095-1187).
__device__ void legetab_recur(int m, double targ, double lambda, double boxdim, double* fint, double* fx,
                              double* fxx) {
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[kMaxOrder + 2], fpp[kMaxOrder + 2];
    double l2 = lambda * lambda;
    double a = (-1 - targ) / lambda, b = (1 - targ) / lambda;
    double ea = exp(-a * a), eb = exp(-b * b);
    double d1 = lambda * (eb - ea), d2 = lambda * (eb + ea);
    fint[0] = sqrtpih * (erf(b) - erf(a));
    if (m > 1) fint[1] = -d1 / 2 + targ * fint[0];
    if (m > 2) fint[2] = 0.75 * (-d2 - targ * d1 + (l2 - 2.0 / 3 + 2 * targ * targ) * fint[0]);
    if (m > 3) fint[3] = 5.0 * (-d1 + 2 * targ * (4.0 * fint[2] / 3 - fint[0] / 3) - (0.4 - 4 * l2) * fint[1]) / 8;
    for (int k = 5; k <= m; k++) {
        int nn = k - 3;
        double v = targ * (fint[k - 2] - fint[k - 4]) +
                   (2 * nn + 1) * (l2 / 2 + 1.0 / ((2 * nn + 3) * (2 * nn - 1))) * fint[k - 3] +
                   (nn - 1.0) * fint[k - 5] / (2 * nn - 1);
        fint[k - 1] = (2 * nn + 3.0) * v / (nn + 2);
    }
    fp[0] = 0;
    if (m > 1) fp[1] = fint[0] * lambda;
    for (int k = 3; k <= m; k++) fp[k - 1] = fp[k - 3] + (2 * k - 3) * fint[k - 2] * lambda;
    fpp[0] = 0;
    if (m > 1) fpp[1] = 0;
    for (int k = 3; k <= m; k++) fpp[k - 1] = fpp[k - 3] + (2 * k - 3) * fp[k - 2];
    int sg = 1;
    for (int k = 1; k <= m; k++) {
        double px1 = (k - 1) * k / 2.0;
        fxx[k - 1] = fpp[k - 1] - px1 * (eb + sg * ea) - 2 * (b * eb - sg * a * ea) / lambda;
        sg = -sg;
    }
    double ww = boxdim * lambda / 2;
    for (int k = 0; k < m; k++) {
        fint[k] *= ww;
        fxx[k] *= 2 / boxdim;
    }
    for (int k = 2; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d2 / lambda;
    for (int k = 1; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d1 / lambda;
}

// Chebyshev counterpart (bfgt/bfgt_volrouts.f:1191-1294).
__device__ void chebtab_recur(int m, double targ, double lambda, double boxdim, double* fint, double* fx,
                              double* fxx) {
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[kMaxOrder + 2], fpp[kMaxOrder + 2];
    double l2 = lambda * lambda;
    double a = (-1 - targ) / lambda, b = (1 - targ) / lambda;
    double ea = exp(-a * a), eb = exp(-b * b);
    double d1 = lambda * (eb - ea), d2 = lambda * (eb + ea);
    fint[0] = sqrtpih * (erf(b) - erf(a));
    if (m > 1) fint[1] = -d1 / 2 + targ * fint[0];
    if (m > 2) fint[2] = -d2 - targ * d1 + (l2 - 1 + 2 * targ * targ) * fint[0];
    if (m > 3) fint[3] = -d1 + 2 * targ * fint[2] - (1 - 4 * l2) * fint[1];
    int sg = 1;
    for (int k = 5; k <= m; k++) {
        double dc = (sg == 1) ? d2 : d1;
        int nn = k - 1;
        double dd = (nn - 1.0) / (nn - 3);
        fint[k - 1] = 2 * targ * fint[k - 2] + 2 * ((nn - 1) * l2 + 1.0 / (nn - 3)) * fint[k - 3] +
                      dd * (fint[k - 5] - 2 * targ * fint[k - 4]) - (1 - dd) * dc;
        sg = -sg;
    }
    fp[0] = 0;
    if (m > 1) fp[1] = fint[0] * lambda;
    if (m > 2) fp[2] = 4 * fint[1] * lambda;
    for (int k = 4; k <= m; k++) fp[k - 1] = (fp[k - 3] / (k - 3) + 2 * fint[k - 2] * lambda) * (k - 1);
    fpp[0] = 0;
    if (m > 1) fpp[1] = 0;
    if (m > 2) fpp[2] = 4 * fint[0] * lambda;
    for (int k = 4; k <= m; k++) fpp[k - 1] = (fpp[k - 3] / (k - 3) + 2 * fp[k - 2]) * (k - 1);
    sg = 1;
    for (int k = 1; k <= m; k++) {
        double tp1 = (double)(k - 1) * (k - 1);
        fxx[k - 1] = fpp[k - 1] - tp1 * (eb + sg * ea) - 2 * (b * eb - sg * a * ea) / lambda;
        sg = -sg;
    }
    double ww = boxdim * lambda / 2, sc = 2 / boxdim;
    for (int k = 0; k < m; k++) {
        fint[k] *= ww;
        fxx[k] *= sc;
    }
    for (int k = 2; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d2 / lambda;
    for (int k = 1; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d1 / lambda;
}

// sparsity pattern of one n x n column-major table (element (m, j) at m + n*j):
// ind[2*j], ind[2*j+1] first/last row above eps*max/n (1-based, 0/-1 when
// empty) for column j, slot n the first/last non-empty column
// (bfgt/bfgt_volrouts.f:1300-1378).  Runs on one thread.
__device__ void sparse_pattern(double eps, int n, const double* a, int* ind) {
    for (int i = 0; i <= n; i++) {
        ind[2 * i] = 0;
        ind[2 * i + 1] = -1;
    }
    double dmax = 0;
    for (int i = 0; i < n * n; i++) dmax = fmax(dmax, fabs(a[i]));
    if (dmax <= 2e-16) return;
    double thr = eps * dmax / n;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++)
            if (fabs(a[j + n * i]) > thr) {
                ind[2 * i] = j + 1;
                break;
            }
        for (int j = n - 1; j >= 0; j--)
            if (fabs(a[j + n * i]) > thr) {
                ind[2 * i + 1] = j + 1;
                break;
            }
    }
    for (int i = 0; i < n; i++)
        if (ind[2 * i] > 0) {
            ind[2 * n] = i + 1;
            break;
        }
    for (int i = n - 1; i >= 0; i--)
        if (ind[2 * i + 1] > 0) {
            ind[2 * n + 1] = i + 1;
            break;
        }
}

// The offsets mk_loctab_all computes for targets refined up to mref times
// (bfgt/bfgt_volrouts.f:548-563): k(i0, j0) for source/target size ratios
// 2^i0, i0 = mref+1 .. -1, each also giving its mirror -k; the tables of a
// level occupy slots -half..half (half = 2^(mref+1) (mref+3),
// bfgt/boxfgt.f:625), target centre k / 2^(mref+1) source half-widths away.
struct LoctabSet {
    int nblk, half;
    double cdiv;
    int k[64], i0[64];
};

LoctabSet loctab_set(int mref) {
    LoctabSet S{};
    const int m0 = mref + 1;
    S.half = (1 << (mref + 1)) * (mref + 3);
    S.cdiv = (double)(1 << m0);
    for (int i0 = m0; i0 >= -1; i0--)
        for (int j0 = 1; j0 <= (1 << m0); j0++) {
            int k;
            if (i0 > 0) k = -(1 << (m0 - i0)) - (1 << (m0 - i0 + 1)) * (j0 - 1);
            else if (i0 == 0) k = -(1 << (m0 - i0 + 1)) * (j0 - 1);
            else k = -(1 << m0) - (1 << (m0 + 1)) * (j0 - 1);
            if (k < -S.half) break;
            S.k[S.nblk] = k;
            S.i0[S.nblk] = i0;
            S.nblk++;
        }
    return S;
}

// Near-field tables.  grid (set.nblk, nlevels+1): blockIdx.x picks one of the
// set's computed offsets k (for mref = 0 the six k in {-1,-3, 0,-4, -2,-6}:
// source twice, equal, half the target size); the block also writes the
// mirror offset -k.  block 256.
__global__ void k_loctab(double eps, int ipoly, int n, double delta, const double* boxsize, const double* nodes,
                         LoctabSet set, double* tloc, int* indloc) {
    NodeLayout L{n};
    int k = set.k[blockIdx.x];
    int i0 = set.i0[blockIdx.x];
    const int nslot = 2 * set.half + 1;
    int lev = blockIdx.y;
    double boxdim = boxsize[lev];
    extern __shared__ double sh[];
    const int nn = n * n;
    double* t0 = sh;            // raw tables [m + n*j]
    double* t1 = t0 + nn;
    double* t2 = t1 + nn;
    double* A = t2 + nn;        // folded tables, 3 x nn (offset k)
    double* B = A + 3 * nn;     // mirrored tables, 3 x nn (offset -k)
    __shared__ int ind3[3][2 * (kMaxOrder + 1)];
    __shared__ int indm[2][2 * (kMaxOrder + 1)];

    double scale = ldexp(1.0, i0);
    double cen = k / set.cdiv;
    double bs = boxdim * scale;
    double sigma = 4 * delta / (bs * bs);
    double lambda = sqrt(sigma);
    double dmax = sqrt(log(1 / 1e-16) * sigma);
    int tid = threadIdx.x;
    if (lambda <= 0.125) {
        if (tid < n) {
            int j = tid;
            double targ = nodes[L.xq() + j] / scale + cen;
            double dis = 0;
            if (targ < -1) dis = -1 - targ;
            if (targ > 1) dis = targ - 1;
            double fint[kMaxOrder + 2], fx[kMaxOrder + 2], fxx[kMaxOrder + 2];
            if (dis > dmax) {
                for (int m = 0; m < n; m++) t0[m + n * j] = t1[m + n * j] = t2[m + n * j] = 0.0;
            } else {
                if (ipoly == 0) legetab_recur(n, targ, lambda, bs, fint, fx, fxx);
                else chebtab_recur(n, targ, lambda, bs, fint, fx, fxx);
                for (int m = 0; m < n; m++) {
                    t0[m + n * j] = fint[m];
                    t1[m + n * j] = fx[m];
                    t2[m + n * j] = fxx[m];
                }
            }
        }
    } else {
        double bsh = bs / 2;
        for (int e = tid; e < nn; e += blockDim.x) {
            int m = e % n, j = e / n;
            double targ = nodes[L.xq() + j] / scale + cen;
            double dis = 0;
            if (targ < -1) dis = -1 - targ;
            if (targ > 1) dis = targ - 1;
            double r = 0, rx = 0, rxx = 0;
            if (!(dis > dmax)) {
                double pv[kMaxOrder + 2];
                for (int i = 0; i < kNq; i++) {
                    double xi = nodes[L.xn() + i];
                    double dx = targ - xi;
                    double we = (fabs(dx) < dmax) ? exp(-dx * dx / sigma) * nodes[L.wn() + i] : 0.0;
                    if (ipoly == 0) legendre_pols(xi, m, pv);
                    else cheb_pols(xi, m, pv);
                    double dd = -2 * dx / sigma * 2 / bs;
                    double pw = pv[m] * we;
                    r += pw;
                    rx += pw * dd;
                    rxx += pw * (dd * dd - 2 / delta);
                }
                r *= bsh;
                rx *= bsh;
                rxx *= bsh;
            }
            t0[m + n * j] = r;
            t1[m + n * j] = rx;
            t2[m + n * j] = rxx;
        }
    }
    __syncthreads();
    // fold U: A_t(i, j) = sum_l U(l, i) t_t(l, j)
    for (int e = tid; e < nn; e += blockDim.x) {
        int i = e % n, j = e / n;
        double s0 = 0, s1 = 0, s2 = 0;
        for (int l = 0; l < n; l++) {
            double u = nodes[L.U() + l * n + i];
            s0 += u * t0[l + n * j];
            s1 += u * t1[l + n * j];
            s2 += u * t2[l + n * j];
        }
        A[e] = s0;
        A[nn + e] = s1;
        A[2 * nn + e] = s2;
    }
    __syncthreads();
    // mirror: B(m, j) = A(n-1-m, n-1-j), derivative table negated.  For
    // k = 0 the reference performs this in place, which leaves element L
    // equal to the original element max(L, n*n-1-L) (negated for the
    // derivative table in the lower half and at the centre).
    for (int e = tid; e < nn; e += blockDim.x) {
        int r = nn - 1 - e;
        if (k != 0) {
            B[e] = A[r];
            B[nn + e] = -A[nn + r];
            B[2 * nn + e] = A[2 * nn + r];
        } else {
            int src = e > r ? e : r;
            B[e] = A[src];
            B[nn + e] = (e <= r) ? -A[nn + r] : A[nn + e];
            B[2 * nn + e] = A[2 * nn + src];
        }
    }
    __syncthreads();
    // sparsity of the mirrored (slot -k) and original (slot k) tables
    for (int s = 0; s < 2; s++) {
        const double* T = (s == 0) ? B : A;
        if (k == 0) T = B;  // slot 0 holds the symmetrised table only
        if (tid < 3) sparse_pattern(eps, n, T + tid * nn, ind3[tid]);
        __syncthreads();
        if (tid <= n) {
            int j = tid;
            int i2 = ind3[0][2 * j + 1];
            if (i2 < ind3[1][2 * j + 1]) i2 = ind3[1][2 * j + 1];
            if (i2 < ind3[2][2 * j + 1]) i2 = ind3[2][2 * j + 1];
            int i1 = ind3[0][2 * j];
            if (ind3[1][2 * j] != 0 && i1 > ind3[1][2 * j]) i1 = ind3[1][2 * j];
            if (ind3[2][2 * j] != 0 && i1 > ind3[2][2 * j]) i1 = ind3[2][2 * j];
            if (i2 > 0 && i1 == 0) i1 = 1;
            indm[s][2 * j] = i1;
            indm[s][2 * j + 1] = i2;
        }
        __syncthreads();
    }
    // masked writes: tloc[lev][q][t][j][m]
    for (int s = 0; s < 2; s++) {
        int q = ((s == 0) ? -k : k) + set.half;
        if (k == 0 && s == 1) break;
        const double* T = (s == 0 || k == 0) ? B : A;
        const int* im = indm[s];
        int* out_ind = indloc + ((size_t)lev * nslot + q) * 2 * (n + 1);
        for (int e = tid; e < 2 * (n + 1); e += blockDim.x) out_ind[e] = im[e];
        double* out = tloc + ((size_t)lev * nslot + q) * 3 * nn;
        int jr0 = im[2 * n], jr1 = im[2 * n + 1];
        for (int e = tid; e < 3 * nn; e += blockDim.x) {
            int t = e / nn, r = e % nn;
            int j = r / n, m = r % n;  // output layout [t][j][m]
            bool keep = (j + 1 >= jr0) && (j + 1 <= jr1) && (m + 1 >= im[2 * j]) && (m + 1 <= im[2 * j + 1]);
            out[e] = keep ? T[t * nn + m + n * j] : 0.0;
        }
    }
}

// mp -> loc diagonal shifts for offsets (kx, ky) in {-1,0,1}^2 at the cutoff
// level, index (ky+1)*3 + (kx+1); element j2*H + j1 for x frequency j2 < npw
// and retained y frequency j1 < H (the device expansion layout).
__global__ void k_m2l_shifts(int npw, double xmin, const double* ts, cplx* w) {
    int H = (npw + 1) / 2;
    int nexp = npw * H;
    int k = blockIdx.x;
    int kx = k % 3 - 1, ky = k / 3 - 1;
    for (int e = threadIdx.x; e < nexp; e += blockDim.x) {
        int j2 = e / H, j1 = e % H;
        double sx, cx, sy, cy;
        sincos(ts[j2] * xmin, &sx, &cx);
        sincos(ts[j1] * xmin, &sy, &cy);
        double axr = kx == 0 ? 1.0 : cx, axi = kx == 0 ? 0.0 : kx * sx;
        double ayr = ky == 0 ? 1.0 : cy, ayi = ky == 0 ? 0.0 : ky * sy;
        w[(size_t)k * nexp + e] = cplx{axr * ayr - axi * ayi, axr * ayi + axi * ayr};
    }
}

}  // namespace

int build_phase_tables(Plan& p);

int build_tables(Plan& p) {
    hipEvent_t& e0 = p.ev[8];
    hipEvent_t& e1 = p.ev[9];
    BFGT_HIP_CHECK(hipEventRecord(e0, p.stream));
    const int n = p.n;
    k_node_rules<<<1, 64, 0, p.stream>>>(n, p.ipoly, p.d_nodes);
    BFGT_HIP_CHECK(hipGetLastError());
    // near-field tables for every level
    size_t sh = sizeof(double) * 9 * n * n;
    if (sh > 65536) BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)k_loctab, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    const LoctabSet set0 = loctab_set(0);
    k_loctab<<<dim3(set0.nblk, p.nlevels + 1), 256, sh, p.stream>>>(p.eps, p.ipoly, n, p.delta, p.d_boxsize,
                                                                    p.d_nodes, set0, p.d_tloc, p.d_indloc);
    BFGT_HIP_CHECK(hipGetLastError());
    if (p.has_pw) {
        int nlev_pw = p.nlevels - p.nlevstart + 1;
        k_poly2pw<<<dim3(p.npw - p.npw / 2, nlev_pw), 64, 0, p.stream>>>(
            n, p.ipoly, p.npw, p.nlevstart, p.d_ws, p.d_ts, p.delta0, p.d_boxsize, p.d_nodes, p.d_tpoly);
        BFGT_HIP_CHECK(hipGetLastError());
        k_pw2pgh<<<nlev_pw, 256, 0, p.stream>>>(n, p.npw, p.nlevstart, p.d_ts, p.d_boxsize, p.delta0, p.d_nodes,
                                                p.d_tpw);
        BFGT_HIP_CHECK(hipGetLastError());
        double xmin = p.boxsize[p.nlevstart] / std::sqrt(p.delta0);
        k_m2l_shifts<<<9, 256, 0, p.stream>>>(p.npw, xmin, p.d_ts, p.d_wshift);
        BFGT_HIP_CHECK(hipGetLastError());
        int rc = build_phase_tables(p);
        if (rc) return rc;
    }
    size_t nind = (size_t)(p.nlevels + 1) * 13 * 2 * (n + 1);
    p.h_indloc.resize(nind);
    BFGT_HIP_CHECK(hipMemcpyAsync(p.h_indloc.data(), p.d_indloc, nind * sizeof(int), hipMemcpyDeviceToHost,
                                  p.stream));
    BFGT_HIP_CHECK(hipEventRecord(e1, p.stream));
    return BFGT_OK;
}

// Near-field tables for targets refined up to mref times below the tree's
// leaves, every level: tloc [level][2 half + 1][3][target][source], masked
// like the plan's own, slots the enumeration does not reach left zero.
int build_loctab_refined(Plan& p, int mref, double* tloc, int* indloc, int* half) {
    const int n = p.n;
    const LoctabSet set = loctab_set(mref);
    const size_t nslot = 2 * set.half + 1;
    BFGT_HIP_CHECK(hipMemsetAsync(tloc, 0, sizeof(double) * (p.nlevels + 1) * nslot * 3 * n * n, p.stream));
    BFGT_HIP_CHECK(hipMemsetAsync(indloc, 0, sizeof(int) * (p.nlevels + 1) * nslot * 2 * (n + 1), p.stream));
    size_t sh = sizeof(double) * 9 * n * n;
    if (sh > 65536) BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)k_loctab, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    k_loctab<<<dim3(set.nblk, p.nlevels + 1), 256, sh, p.stream>>>(p.eps, p.ipoly, n, p.delta, p.d_boxsize,
                                                                   p.d_nodes, set, tloc, indloc);
    BFGT_HIP_CHECK(hipGetLastError());
    *half = set.half;
    return BFGT_OK;
}

// waits for the tables launched by build_tables (the host copy of the
// sparsity bounds included) and records their device time
int finish_tables(Plan& p) {
    BFGT_HIP_CHECK(hipEventSynchronize(p.ev[9]));
    float ms = 0;
    BFGT_HIP_CHECK(hipEventElapsedTime(&ms, p.ev[8], p.ev[9]));
    p.t_tables = ms * 1e-3;
    return BFGT_OK;
}

size_t node_buffer_doubles(int n) { return (size_t)NodeLayout{n}.size(); }

}  // namespace bfgt

namespace bfgt {
namespace {

// PW nodes and weights: midpoint rule of the free-space Gaussian
// (get_pwnodes itype=0, common/fgtterms.f:151-179) or the Fourier series of
// the periodic Gaussian (get_periodic_pwnodes, common/fgtterms.f:189-209).
__global__ void k_pwnodes(int npw, int periodic, double pmax, double bs0, double delta, double* ws, double* ts) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npw) return;
    int npw2 = npw / 2;
    if (!periodic) {
        double h = pmax / npw2;
        double w = h / (2.0 * sqrt(M_PI));
        double t = (j - npw2 + 0.5) * h;
        ts[j] = t;
        ws[j] = w * exp(-t * t / 4);
    } else {
        double h = 2 * M_PI / bs0;
        double w = sqrt(M_PI * delta) / bs0;
        double hw = h * sqrt(delta) / 2;
        int jj = j - npw2;
        ts[j] = jj * h;
        ws[j] = w * exp(-(hw * jj) * (hw * jj));
    }
}

}  // namespace

int build_pw_nodes(Plan& p) {
    bool periodic = (p.iperiod_in == 1 && p.npwlevel == 0);
    k_pwnodes<<<(p.npw + 255) / 256, 256, 0, p.stream>>>(p.npw, periodic ? 1 : 0, p.pmax, p.boxsize[0], p.delta,
                                                         p.d_ws, p.d_ts);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

}  // namespace bfgt

namespace bfgt {
namespace {

// Leaf-to-ancestor phases: for every level l >= L0 and every cell j of a
// level-L0 box at level l, phase[levbase[l] + j*pstride + k] =
// exp(i t_k ((j + 1/2) b_l - b_L0/2) / sqrt(delta0)) -- the product of the
// per-level merge/split diagonals (bfgt/bfgt_pwrouts.f:1645-1703) between a
// leaf and its cutoff-level ancestor, evaluated directly; zero for the
// padding frequencies k >= npw.
__global__ void k_phase_tables(int npw, int pstride, int lev0, int nlev, const double* ts, const double* boxsize,
                               double delta0, const int* levbase, cplx* phase) {
    int l = lev0 + blockIdx.y;
    if (l > nlev) return;
    int nj = 1 << (l - lev0);
    double sq = sqrt(delta0);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nj * pstride; e += gridDim.x * blockDim.x) {
        int j = e / pstride, k = e % pstride;
        cplx v{0.0, 0.0};
        if (k < npw) {
            double off = ((j + 0.5) * boxsize[l] - 0.5 * boxsize[lev0]) / sq;
            double s, c;
            sincos(ts[k] * off, &s, &c);
            v = cplx{c, s};
        }
        phase[levbase[l] + e] = v;
    }
}

// Zero-padded copies of the PW tables for the matrix-core kernels (layouts in
// the Plan declaration); levels above the cutoff are zero.
__global__ void k_pad_tables(int n, int npw, int ps, int lev0, int nlev, const cplx* __restrict__ tpoly,
                             const cplx* __restrict__ tpw, cplx* __restrict__ tup_f, cplx* __restrict__ tup_a,
                             cplx* __restrict__ tdn_f, cplx* __restrict__ tdn_a) {
    const int per = 16 * ps;  // one padded table of one level
    const size_t tot = (size_t)nlev * 8 * per;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(e / (8 * per));
        const int r = (int)(e % (8 * per));
        const int slot = r / per, w = r % per;
        const bool lev_ok = l >= lev0;
        const cplx* T = tpoly + (size_t)l * npw * n;      // [m][k]
        const cplx* E = tpw + (size_t)l * 3 * npw * n;    // [t][k][i]
        cplx v{0.0, 0.0};
        if (slot == 0) {  // tup_f [m][k]
            int m = w / ps, k = w % ps;
            if (lev_ok && m < n && k < npw) v = T[m * npw + k];
            tup_f[(size_t)l * per + w] = v;
        } else if (slot <= 3) {  // tup_a [t][m][k]
            int t = slot - 1, m = w / ps, k = w % ps;
            if (lev_ok && m < n && k < npw) {
                v = E[(t * npw + k) * n + m];
                v.y = -v.y;
            }
            tup_a[((size_t)l * 3 + t) * per + w] = v;
        } else if (slot <= 6) {  // tdn_f [t][k][i]
            int t = slot - 4, k = w / 16, i = w % 16;
            if (lev_ok && k < npw && i < n) v = E[(t * npw + k) * n + i];
            tdn_f[((size_t)l * 3 + t) * per + w] = v;
        } else {  // tdn_a [k][i]
            int k = w / 16, i = w % 16;
            if (lev_ok && k < npw && i < n) {
                v = T[i * npw + k];
                v.y = -v.y;
            }
            tdn_a[(size_t)l * per + w] = v;
        }
    }
}

}  // namespace

int build_phase_tables(Plan& p) {
    int nl = p.nlevels - p.nlevstart + 1;
    k_phase_tables<<<dim3(64, nl), 256, 0, p.stream>>>(p.npw, p.pstride, p.nlevstart, p.nlevels, p.d_ts,
                                                      p.d_boxsize, p.delta0, p.d_levbase, p.d_phase);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int build_padded_tables(Plan& p) {
    int nlev = p.nlevels + 1;
    size_t tot = (size_t)nlev * 8 * 16 * p.pstride;
    int blocks = (int)std::min<size_t>((tot + 255) / 256, 4096);
    k_pad_tables<<<blocks, 256, 0, p.stream>>>(p.n, p.npw, p.pstride, p.nlevstart, nlev, p.d_tpoly, p.d_tpw,
                                               p.d_tup_f, p.d_tup_a, p.d_tdn_f, p.d_tdn_a);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

}  // namespace bfgt
