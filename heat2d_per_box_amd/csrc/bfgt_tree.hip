// Adaptive quadtrees in device memory: building a tree that resolves a user
// function, re-meshing it for a new function with values carried along, and
// level restriction -- the tree side of the heat solvers' time step
// (common/tree_ext.f vol_tree_mem2 :996, vol_tree_build2 :1363,
// vol_tree_adap_fv :2337; common/tree_vol_coeffs_nd.f).
//
// Everything is level-synchronous.  A step evaluates the tail of every
// candidate box's tensor-product coefficients (one workgroup per box), flags
// boxes, numbers the children of flagged boxes by a prefix sum over the
// candidates (so children follow their parents' order, as in the reference's
// builder) and creates them; values on new boxes come from the user callback
// or from the parent's polynomial evaluated on the children's nodes.  Level
// restriction is a closure over cells: every refined cell at level k needs
// the parents of its 3x3 neighbourhood refined at level k-1 (periodic wrap
// or clipping at the domain edge), computed from the finest level up with a
// device hash set of (level, ix, iy) keys and then materialised top-down.
// After every structural change the boxes are renumbered by a breadth-first
// pass over the child links.
//
// Boxes are 0-based here; parents/children are -1 where absent.  The export
// turns them into the reference's 1-based layout.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bfgt_internal.h"

struct bfgt_tree;

namespace bfgt {
namespace {

constexpr int kTreeMaxLevel = 28;          // cell coordinates use 29 bits
constexpr unsigned long long kEmptyKey = ~0ull;

}  // namespace

// --------------------------------------------------------------------------
// 1-D rules on the host: nodes, values -> coefficients, child/parent
// interpolation, tail mask and product weights.

static void poly_values(int ipoly, double x, int n, double* p) {
    p[0] = 1.0;
    if (n == 1) return;
    p[1] = x;
    for (int k = 1; k + 1 < n; k++)
        p[k + 1] = ipoly == 0 ? ((2 * k + 1) * x * p[k] - k * p[k - 1]) / (k + 1) : 2 * x * p[k] - p[k - 1];
}

static void tail_mask(int n, int iptype, std::vector<double>& m, double& rsum) {
    // tens_prod_get_rmask (common/polytens.f:1579-1621): (p1+1)^q + (p2+1)^q >= n^q, q = max(iptype, 1)
    int q = std::max(iptype, 1);
    auto pw = [q](long v) { return q == 1 ? v : v * v; };
    m.assign(n * n, 0.0);
    double cnt = 0;
    for (int ky = 0; ky < n; ky++)
        for (int kx = 0; kx < n; kx++)
            if (pw(kx + 1) + pw(ky + 1) >= pw(n)) {
                m[kx + n * ky] = 1.0;
                cnt += 1;
            }
    rsum = iptype == 2 ? std::sqrt(cnt) : (iptype == 0 ? 1.0 : cnt);
}

TreeRules make_rules(int n, int ipoly, int iptype) {
    TreeRules r;
    r.n = n;
    r.ipoly = ipoly;
    r.iptype = iptype;
    r.x.assign(n, 0.0);
    r.w.assign(n, 0.0);
    std::vector<double> p(n + 1);
    if (ipoly == 0) {
        // Gauss-Legendre by Newton from cosine guesses
        for (int i = 0; i < (n + 1) / 2; i++) {
            double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
            for (int it = 0; it < 100; it++) {
                poly_values(0, z, n + 1, p.data());
                dp = n * (z * p[n] - p[n - 1]) / (z * z - 1.0);
                double dz = p[n] / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16 && it > 2) break;
            }
            poly_values(0, z, n + 1, p.data());
            dp = n * (z * p[n] - p[n - 1]) / (z * z - 1.0);
            if (n % 2 == 1 && i == (n + 1) / 2 - 1) z = 0.0;
            double wt = 2.0 / ((1.0 - z * z) * dp * dp);
            r.x[n - 1 - i] = z;
            r.x[i] = -z;
            r.w[n - 1 - i] = wt;
            r.w[i] = wt;
        }
    } else {
        // Chebyshev nodes cos((2i-1) pi / 2n), ascending, and the weights of
        // chebexps (common/chebexps.f:79-106): (2 - sum_{even j} 4/(j^2-1) T_j(x)) / n
        for (int i = 1; i <= n; i++) r.x[n - i] = std::cos((2 * i - 1) * M_PI / (2 * n));
        for (int i = 0; i < n; i++) {
            double tm2 = 1, tm1 = r.x[i], wt = 2;
            int ic = -1;
            for (int j = 2; j <= n - 1; j++) {
                double tj = 2 * r.x[i] * tm1 - tm2;
                tm2 = tm1;
                tm1 = tj;
                ic = -ic;
                if (ic < 0) continue;
                double rint = -2 * (1.0 / (j + 1) - 1.0 / (j - 1));
                wt -= rint * tj;
            }
            r.w[i] = wt / n;
        }
    }
    r.U.assign(n * n, 0.0);
    for (int j = 0; j < n; j++) {
        poly_values(ipoly, r.x[j], n, p.data());
        for (int k = 0; k < n; k++)
            r.U[k * n + j] = ipoly == 0 ? p[k] * r.w[j] * (2 * k + 1) / 2.0 : (k == 0 ? 1.0 : 2.0) * p[k] / n;
    }
    r.P2C.assign(2 * n * n, 0.0);
    for (int i = 0; i < n; i++) {
        poly_values(ipoly, r.x[i] / 2 - 0.5, n, p.data());
        for (int k = 0; k < n; k++) r.P2C[i * n + k] = p[k];
        poly_values(ipoly, r.x[i] / 2 + 0.5, n, p.data());
        for (int k = 0; k < n; k++) r.P2C[n * n + i * n + k] = p[k];
    }
    int n2 = n / 2;
    r.C2P.assign(2 * std::max(n2, 1) * n, 0.0);
    for (int i = 0; i < n2; i++) {
        poly_values(ipoly, r.x[i] * 2 + 1.0, n, p.data());
        for (int k = 0; k < n; k++) r.C2P[i * n + k] = p[k];
        poly_values(ipoly, r.x[n2 + i] * 2 - 1.0, n, p.data());
        for (int k = 0; k < n; k++) r.C2P[n2 * n + i * n + k] = p[k];
    }
    tail_mask(n, iptype, r.mask, r.rsum);
    tail_mask(n, 2, r.mask2, r.rsum2);
    r.wq.assign(n * n, 0.0);
    for (int iy = 0; iy < n; iy++)
        for (int ix = 0; ix < n; ix++) r.wq[ix + n * iy] = r.w[ix] * r.w[iy];
    return r;
}

namespace {

// --------------------------------------------------------------------------
// device side

struct TreeArrays {
    int *lev = nullptr, *par = nullptr, *nch = nullptr, *ch = nullptr, *cx = nullptr, *cy = nullptr;
    double* ctr = nullptr;
    double* vals = nullptr;
};

struct Rng {  // a contiguous range or an explicit list of boxes
    int m;
    int first;
    const int* list;
    __device__ int box(int i) const { return list ? list[i] : first + i; }
};

__device__ inline unsigned long long cell_key(int l, int x, int y) {
    return ((unsigned long long)l << 58) | ((unsigned long long)(unsigned)x << 29) | (unsigned long long)(unsigned)y;
}

__device__ inline unsigned key_hash(unsigned long long k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return (unsigned)k;
}

// Insert; true when the key was absent.  The table is sized so that it
// never fills (the host bounds the number of keys).
__device__ bool set_insert(unsigned long long* keys, unsigned mask, unsigned long long k) {
    unsigned h = key_hash(k) & mask;
    while (true) {
        unsigned long long prev = atomicCAS(&keys[h], kEmptyKey, k);
        if (prev == kEmptyKey) return true;
        if (prev == k) return false;
        h = (h + 1) & mask;
    }
}

__device__ bool set_find(const unsigned long long* keys, unsigned mask, unsigned long long k) {
    unsigned h = key_hash(k) & mask;
    while (true) {
        unsigned long long v = keys[h];
        if (v == k) return true;
        if (v == kEmptyKey) return false;
        h = (h + 1) & mask;
    }
}

template <typename T>
__global__ void k_fill(T* a, long n, T v) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Node coordinates of the boxes of a range: xy[2*(i*npbox + p) + {0,1}] =
// centre + (x_node/2) * boxsize (x node index fastest, as mesh2d).
__global__ void k_box_points(Rng r, int n, const int* lev, const double* ctr, const double* bsz, const double* xq,
                             double* xy) {
    long npbox = (long)n * n;
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)r.m * npbox) return;
    int i = (int)(t / npbox), p = (int)(t % npbox);
    int b = r.box(i);
    double bs = bsz[lev[b]];
    double gx = xq[p % n] * 0.5, gy = xq[p / n] * 0.5;
    xy[2 * t] = gx * bs + ctr[2 * b];
    xy[2 * t + 1] = gy * bs + ctr[2 * b + 1];
}

// Sampled values (nd, npbox) per box of a list, compact -> into vals.
__global__ void k_scatter_vals(Rng r, int per_box, const double* src, double* vals) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)r.m * per_box) return;
    int i = (int)(t / per_box), q = (int)(t % per_box);
    vals[(long)r.box(i) * per_box + q] = src[t];
}

// Coefficients of one density of one box's values into c[n*n] ([kx + n*ky]);
// v and t are LDS scratch of n*n.  Block-wide; ends synchronised.
__device__ void box_coefs(int n, int nd, int d, const double* vals_b, const double* U, double* v, double* t,
                          double* c) {
    int nn = n * n;
    for (int p = threadIdx.x; p < nn; p += blockDim.x) v[p] = vals_b[(long)p * nd + d];
    __syncthreads();
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        int kx = q % n, iy = q / n;
        double s = 0.0;
        for (int ix = 0; ix < n; ix++) s += U[kx * n + ix] * v[ix + n * iy];
        t[q] = s;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        int kx = q % n, ky = q / n;
        double s = 0.0;
        for (int iy = 0; iy < n; iy++) s += U[ky * n + iy] * t[kx + n * iy];
        c[q] = s;
    }
    __syncthreads();
}

// Tail norm of coefficients c under mask (iptype 0 max, 1 sum, 2 sqrt of sum
// of squares); block-wide reduction through red[blockDim.x].
__device__ double tail_norm(int nn, int iptype, const double* c, const double* mask, double* red) {
    double a = 0.0;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        if (mask[q] > 0.5) {
            double v = c[q];
            if (iptype == 0)
                a = fmax(a, fabs(v));
            else if (iptype == 1)
                a += fabs(v);
            else
                a += v * v;
        }
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = iptype == 0 ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                                            : red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return iptype == 2 ? sqrt(r) : r;
}

__device__ inline double pow_eta(double x, double eta) {
    if (eta == 1.0) return x;
    if (eta == 0.0) return 1.0;
    return pow(x, eta);
}

constexpr int kTB = 64;  // threads of the per-box kernels

// Split flags: flag[b] = 1 when box b's values are not resolved:
// max_d tail(coefs_d) * (boxsize/2)^eta / rsum > thr[level]; with big_split,
// boxes larger than 5 are split regardless (vol_tree_find_box_refine).
__global__ __launch_bounds__(kTB) void k_unresolved(Rng r, int n, int nd, const double* vals, const double* Ug,
                                                    const double* maskg, double rsum, int iptype, double eta,
                                                    const int* lev, const double* bsz, const double* thr,
                                                    int big_split, int* flag) {
    extern __shared__ double sh[];
    int nn = n * n;
    double *U = sh, *mask = U + nn, *v = mask + nn, *t = v + nn, *c = t + nn, *red = c + nn;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        U[q] = Ug[q];
        mask[q] = maskg[q];
    }
    __syncthreads();
    int b = r.box(blockIdx.x);
    int l = lev[b];
    double e = 0.0;
    for (int d = 0; d < nd; d++) {
        box_coefs(n, nd, d, vals + (long)b * nn * nd, U, v, t, c);
        double a = tail_norm(nn, iptype, c, mask, red);
        if (a > e) e = a;
    }
    if (threadIdx.x == 0) {
        double err = e * pow_eta(bsz[l] / 2, eta);
        err = err / rsum;
        int f = err > thr[l];
        if (big_split && bsz[l] > 5.0) f = 1;
        flag[b] = f;
    }
}

// take[i] = flag[b] && b is a leaf, for the boxes of a range
__global__ void k_take(Rng r, const int* flag, const int* nch, int* take) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.m) return;
    int b = r.box(i);
    take[i] = (flag[b] && nch[b] == 0) ? 1 : 0;
}

// Create the children of the taken boxes of a range: the k-th taken box
// (k = off[i]) gets children base + 4k .. base + 4k + 3 and is recorded in
// parents[pbase + k].  Child centres: parent + sign * boxsize(l+1)/2.
__global__ void k_split(Rng r, const int* take, const int* off, int base, int pbase, TreeArrays a, const double* bsz,
                        int* parents) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.m || !take[i]) return;
    int b = r.box(i);
    int k = off[i];
    int c0 = base + 4 * k;
    parents[pbase + k] = b;
    int l = a.lev[b] + 1;
    double h = bsz[l] / 2;
    a.nch[b] = 4;
    for (int j = 0; j < 4; j++) {
        int c = c0 + j;
        a.ch[4 * b + j] = c;
        a.lev[c] = l;
        a.par[c] = b;
        a.nch[c] = 0;
        for (int q = 0; q < 4; q++) a.ch[4 * c + q] = -1;
        a.cx[c] = 2 * a.cx[b] + (j & 1);
        a.cy[c] = 2 * a.cy[b] + (j >> 1);
        a.ctr[2 * c] = a.ctr[2 * b] + ((j & 1) ? h : -h);
        a.ctr[2 * c + 1] = a.ctr[2 * b + 1] + ((j >> 1) ? h : -h);
    }
}

// Values of the four children of each listed parent: the parent's
// polynomial (coefficients from its values) at the children's nodes.
__global__ __launch_bounds__(kTB) void k_to_children(int m, const int* parents, int n, int nd, const int* ch,
                                                     double* vals, const double* Ug, const double* P2Cg) {
    extern __shared__ double sh[];
    int nn = n * n;
    double *U = sh, *P = U + nn, *v = P + 2 * nn, *t = v + nn, *c = t + nn;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) U[q] = Ug[q];
    for (int q = threadIdx.x; q < 2 * nn; q += blockDim.x) P[q] = P2Cg[q];
    __syncthreads();
    int b = parents[blockIdx.x];
    for (int d = 0; d < nd; d++) {
        box_coefs(n, nd, d, vals + (long)b * nn * nd, U, v, t, c);
        for (int j = 0; j < 4; j++) {
            const double* Px = P + (j & 1) * nn;
            const double* Py = P + (j >> 1) * nn;
            // t[ix + n*ky] = sum_kx Px[ix][kx] c[kx + n*ky]
            for (int q = threadIdx.x; q < nn; q += blockDim.x) {
                int ix = q % n, ky = q / n;
                double s = 0.0;
                for (int kx = 0; kx < n; kx++) s += Px[ix * n + kx] * c[kx + n * ky];
                t[q] = s;
            }
            __syncthreads();
            double* out = vals + (long)ch[4 * b + j] * nn * nd;
            for (int q = threadIdx.x; q < nn; q += blockDim.x) {
                int ix = q % n, iy = q / n;
                double s = 0.0;
                for (int ky = 0; ky < n; ky++) s += Py[iy * n + ky] * t[ix + n * ky];
                out[(long)q * nd + d] = s;
            }
            __syncthreads();
        }
    }
}

// Merging test for the boxes of one level whose four children are leaves
// (vol_tree_coarsen_f2): the children's polynomials of fun (fv) and of the
// carried values (uv) evaluated on the parent's nodes in each quadrant are
// written into the parent; when both l2 tails (eta = 0) over rsum2 are at
// most reps the children are dropped.
__global__ __launch_bounds__(kTB) void k_merge_children(Rng r, int n, int nd, TreeArrays a, double* fv,
                                                        const double* Ug, const double* C2Pg, const double* mask2g,
                                                        double rsum2, double reps) {
    extern __shared__ double sh[];
    int nn = n * n, n2 = n / 2;
    double *U = sh, *P = U + nn, *mask = P + 2 * n2 * n, *v = mask + nn, *t = v + nn, *c = t + nn, *red = c + nn;
    int b = r.box(blockIdx.x);
    if (a.nch[b] != 4) return;
    for (int j = 0; j < 4; j++)
        if (a.nch[a.ch[4 * b + j]] != 0) return;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        U[q] = Ug[q];
        mask[q] = mask2g[q];
    }
    for (int q = threadIdx.x; q < 2 * n2 * n; q += blockDim.x) P[q] = C2Pg[q];
    __syncthreads();
    double err[2] = {0.0, 0.0};
    for (int arr = 0; arr < 2; arr++) {
        double* vals = arr == 0 ? fv : a.vals;
        for (int d = 0; d < nd; d++) {
            for (int j = 0; j < 4; j++) {
                int cj = a.ch[4 * b + j];
                box_coefs(n, nd, d, vals + (long)cj * nn * nd, U, v, t, c);
                const double* Px = P + (j & 1) * n2 * n;
                const double* Py = P + (j >> 1) * n2 * n;
                // t[i + n2*ky] = sum_kx Px[i][kx] c[kx + n*ky], i < n2
                for (int q = threadIdx.x; q < n2 * n; q += blockDim.x) {
                    int i = q % n2, ky = q / n2;
                    double s = 0.0;
                    for (int kx = 0; kx < n; kx++) s += Px[i * n + kx] * c[kx + n * ky];
                    t[q] = s;
                }
                __syncthreads();
                double* out = vals + (long)b * nn * nd;
                int ox = (j & 1) * n2, oy = (j >> 1) * n2;
                for (int q = threadIdx.x; q < n2 * n2; q += blockDim.x) {
                    int i = q % n2, i2 = q / n2;
                    double s = 0.0;
                    for (int ky = 0; ky < n; ky++) s += Py[i2 * n + ky] * t[i + n2 * ky];
                    out[(long)((ox + i) + n * (oy + i2)) * nd + d] = s;
                }
                __syncthreads();
            }
            box_coefs(n, nd, d, vals + (long)b * nn * nd, U, v, t, c);
            double e = tail_norm(nn, 2, c, mask, red);
            if (e > err[arr]) err[arr] = e;
        }
    }
    if (threadIdx.x == 0) {
        double ea = err[0] / rsum2, eb = err[1] / rsum2;
        if (ea <= reps && eb <= reps) {
            for (int j = 0; j < 4; j++) {
                int cj = a.ch[4 * b + j];
                a.par[cj] = -1;
                a.ch[4 * b + j] = -1;
            }
            a.nch[b] = 0;
        }
    }
}

// Per-box integral of the values for the norm estimate (update_rints,
// common/tree_vol_coeffs_nd.f:945-1070), in the reference's loop order:
// iptype 0 max |f|, 1 sum |f| w rsc, 2 sum f^2 w rsc.
__global__ void k_box_integral(Rng r, int npbox, int nd, const double* vals, const double* wq, double rsc, int iptype,
                               double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.m) return;
    int b = r.box(i);
    const double* f = vals + (long)b * npbox * nd;
    double s = 0.0;
    for (int p = 0; p < npbox; p++)
        for (int d = 0; d < nd; d++) {
            double v = f[(long)p * nd + d];
            if (iptype == 0)
                s = fmax(s, fabs(v));
            else if (iptype == 1)
                s = __dadd_rn(s, __dmul_rn(__dmul_rn(fabs(v), wq[p]), rsc));
            else
                s = __dadd_rn(s, __dmul_rn(__dmul_rn(__dmul_rn(v, v), wq[p]), rsc));
        }
    out[i] = s;
}

// ---- level restriction -----------------------------------------------------

// refined cells of one level into the set
__global__ void k_refined_cells(Rng r, const int* nch, const int* lev, const int* cx, const int* cy,
                                unsigned long long* keys, unsigned mask, int2* list, int* count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.m) return;
    int b = r.box(i);
    if (nch[b] == 0) return;
    set_insert(keys, mask, cell_key(lev[b], cx[b], cy[b]));
    int k = atomicAdd(count, 1);
    list[k] = make_int2(cx[b], cy[b]);
}

// For the refined cells of level k: the parents of their 3x3 neighbourhoods
// must be refined at level k-1; new ones are appended to out.
__global__ void k_close(int m, const int2* cells, int k, int iperiod, unsigned long long* keys, unsigned mask,
                        int2* out, int* count, int out_cap) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int N = 1 << k;
    int2 c = cells[i];
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            int qx = c.x + dx, qy = c.y + dy;
            if (iperiod) {
                qx = (qx + N) & (N - 1);
                qy = (qy + N) & (N - 1);
            } else if (qx < 0 || qy < 0 || qx >= N || qy >= N) {
                continue;
            }
            int px = qx >> 1, py = qy >> 1;
            if (set_insert(keys, mask, cell_key(k - 1, px, py))) {
                int j = atomicAdd(count, 1);
                if (j < out_cap) out[j] = make_int2(px, py);
            }
        }
}

// flag[b] = leaf b's cell is in the set (it must be split)
__global__ void k_in_set(Rng r, const int* nch, const int* lev, const int* cx, const int* cy,
                         const unsigned long long* keys, unsigned mask, int* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.m) return;
    int b = r.box(i);
    flag[b] = (nch[b] == 0 && set_find(keys, mask, cell_key(lev[b], cx[b], cy[b]))) ? 1 : 0;
}

// ---- renumbering and colleagues ----------------------------------------------

__global__ void k_level_ids(int m, const int* cur, int base, int* newid, const int* nch, int* cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    newid[cur[i]] = base + i;
    cnt[i] = nch[cur[i]] ? 4 : 0;
}

__global__ void k_level_next(int m, const int* cur, const int* off, const int* ch, int* next) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int b = cur[i];
    if (ch[4 * b] < 0) return;
    for (int j = 0; j < 4; j++) next[off[i] + j] = ch[4 * b + j];
}

__global__ void k_gather(int nold, const int* newid, int per_box, TreeArrays s, TreeArrays d, const double* fvs,
                         double* fvd) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int b = (int)(t / kTB);
    if (b >= nold) return;
    int nb = newid[b];
    if (nb < 0) return;
    int lane = (int)(t % kTB);
    if (lane == 0) {
        d.lev[nb] = s.lev[b];
        d.par[nb] = s.par[b] >= 0 ? newid[s.par[b]] : -1;
        d.nch[nb] = s.nch[b];
        for (int j = 0; j < 4; j++) d.ch[4 * nb + j] = s.ch[4 * b + j] >= 0 ? newid[s.ch[4 * b + j]] : -1;
        d.cx[nb] = s.cx[b];
        d.cy[nb] = s.cy[b];
        d.ctr[2 * nb] = s.ctr[2 * b];
        d.ctr[2 * nb + 1] = s.ctr[2 * b + 1];
    }
    for (int q = lane; q < per_box; q += kTB) {
        d.vals[(long)nb * per_box + q] = s.vals[(long)b * per_box + q];
        if (fvs) fvd[(long)nb * per_box + q] = fvs[(long)b * per_box + q];
    }
}

// Colleagues of the boxes of one level: the children of the parent's
// colleagues, in that order, that touch the box (computecoll,
// common/tree_routs_nd.f:106-242; periodic wrap for iperiod = 1).
__global__ void k_colleagues(int first, int m, int l, int iperiod, const int* par, const int* ch, const int* cx,
                             const int* cy, int* ncoll, int* coll) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int b = first + i;
    int N = 1 << l;
    int p = par[b], cnt = 0;
    for (int q = 0; q < ncoll[p]; q++) {
        int cp = coll[9 * p + q];
        for (int j = 0; j < 4; j++) {
            int k = ch[4 * cp + j];
            if (k < 0) continue;
            int dx = abs(cx[k] - cx[b]), dy = abs(cy[k] - cy[b]);
            if (iperiod) {
                dx = min(dx, N - dx);
                dy = min(dy, N - dy);
            }
            if (dx <= 1 && dy <= 1 && cnt < 9) coll[9 * b + cnt++] = k;
        }
    }
    ncoll[b] = cnt;
    for (int q = cnt; q < 9; q++) coll[9 * b + q] = -1;
}

// Export in the reference layout (1-based box numbers).
__global__ void k_export(int nb, TreeArrays a, const int* ncoll, const int* coll, int* ilevel, int* iparent,
                         int* nchild, int* ichild, int* ncollo, int* collo, double* centers) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    ilevel[b] = a.lev[b];
    iparent[b] = a.par[b] >= 0 ? a.par[b] + 1 : -1;
    nchild[b] = a.nch[b];
    for (int j = 0; j < 4; j++) ichild[4 * b + j] = a.ch[4 * b + j] >= 0 ? a.ch[4 * b + j] + 1 : -1;
    ncollo[b] = ncoll[b];
    for (int j = 0; j < 9; j++) collo[9 * b + j] = coll[9 * b + j] >= 0 ? coll[9 * b + j] + 1 : -1;
    centers[2 * b] = a.ctr[2 * b];
    centers[2 * b + 1] = a.ctr[2 * b + 1];
}

__global__ void k_leaf_flags(int nb, const int* nch, int* f) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) f[b] = nch[b] == 0 ? 1 : 0;
}

__global__ void k_compact(int nb, const int* f, const int* off, int* list) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb && f[b]) list[off[b]] = b;
}

inline int blocks(long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

// --------------------------------------------------------------------------
// the tree object

struct DeviceTree {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int nd = 1, n = 8, npbox = 64, ipoly = 0, iperiod = 1, iptype = 2;
    double eta = 1.0, boxlen = 1.0;
    int cap = 0, nboxes = 0, nlevels = 0;
    std::vector<int> lfirst, lcount;  // per level (0-based first box, count)
    std::vector<double> boxsize;      // 0..kTreeMaxLevel
    std::vector<double> thr_host;     // split thresholds per level (source of an asynchronous copy)
    TreeRules rules;

    TreeArrays A, B;  // current arrays and the renumbering target
    int *ncoll = nullptr, *coll = nullptr;
    double* fv = nullptr;   // user function values during a re-mesh
    double* fvB = nullptr;  // renumbering target of fv
    int *flag = nullptr, *take = nullptr, *off = nullptr, *parents = nullptr, *newid = nullptr, *lista = nullptr,
        *listb = nullptr, *counter = nullptr;
    int2 *cells_a = nullptr, *cells_b = nullptr;
    unsigned long long* keys = nullptr;
    unsigned key_mask = 0;
    double *d_bsz = nullptr, *d_thr = nullptr, *d_xq = nullptr, *d_U = nullptr, *d_P2C = nullptr, *d_C2P = nullptr,
           *d_mask = nullptr, *d_mask2 = nullptr, *d_wq = nullptr, *d_red = nullptr;
    double *d_xy = nullptr, *d_smp = nullptr;
    size_t xy_cap = 0;  // points
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    std::vector<void*> owned;

    ~DeviceTree() {
        for (void* q : owned) (void)hipFree(q);
        if (d_xy) (void)hipFree(d_xy);
        if (d_smp) (void)hipFree(d_smp);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace bfgt

struct bfgt_tree {
    bfgt::DeviceTree t;
};

namespace bfgt {
namespace {

template <typename T>
int dalloc(DeviceTree& t, T** p, size_t count) {
    BFGT_HIP_CHECK(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
    t.owned.push_back((void*)*p);
    return 0;
}

int sync(DeviceTree& t) {
    BFGT_HIP_CHECK(hipStreamSynchronize(t.stream));
    return 0;
}

int read_int(DeviceTree& t, const int* d, int* h) {
    BFGT_HIP_CHECK(hipMemcpyAsync(h, d, sizeof(int), hipMemcpyDeviceToHost, t.stream));
    return sync(t);
}

int alloc_arrays(DeviceTree& t, TreeArrays& a) {
    size_t c = t.cap;
    if (dalloc(t, &a.lev, c) || dalloc(t, &a.par, c) || dalloc(t, &a.nch, c) || dalloc(t, &a.ch, 4 * c) ||
        dalloc(t, &a.cx, c) || dalloc(t, &a.cy, c) || dalloc(t, &a.ctr, 2 * c) ||
        dalloc(t, &a.vals, c * t.npbox * t.nd))
        return BFGT_ERR_HIP;
    return 0;
}

int tree_alloc(DeviceTree& t) {
    size_t c = t.cap;
    if (alloc_arrays(t, t.A) || alloc_arrays(t, t.B)) return BFGT_ERR_HIP;
    if (dalloc(t, &t.ncoll, c) || dalloc(t, &t.coll, 9 * c) || dalloc(t, &t.flag, c) || dalloc(t, &t.take, c) ||
        dalloc(t, &t.off, c + 1) || dalloc(t, &t.parents, c / 4 + 1) || dalloc(t, &t.newid, c) ||
        dalloc(t, &t.lista, c) || dalloc(t, &t.listb, c) || dalloc(t, &t.counter, 4) || dalloc(t, &t.cells_a, c) ||
        dalloc(t, &t.cells_b, c))
        return BFGT_ERR_HIP;
    // level restriction keys: at most ~8 * cap (see close_levels); load <= 1/2
    size_t ks = 1;
    while (ks < 16 * c) ks <<= 1;
    if (dalloc(t, &t.keys, ks)) return BFGT_ERR_HIP;
    t.key_mask = (unsigned)(ks - 1);
    int nn = t.n * t.n, n2 = std::max(t.n / 2, 1);
    if (dalloc(t, &t.d_bsz, kTreeMaxLevel + 1) || dalloc(t, &t.d_thr, kTreeMaxLevel + 1) ||
        dalloc(t, &t.d_xq, t.n) || dalloc(t, &t.d_U, nn) || dalloc(t, &t.d_P2C, 2 * nn) ||
        dalloc(t, &t.d_C2P, 2 * n2 * t.n) || dalloc(t, &t.d_mask, nn) || dalloc(t, &t.d_mask2, nn) ||
        dalloc(t, &t.d_wq, nn) || dalloc(t, &t.d_red, c))
        return BFGT_ERR_HIP;
    size_t bytes = 0;
    BFGT_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, t.take, t.off, (int)c + 1, t.stream));
    t.scan_bytes = bytes;
    if (dalloc(t, (char**)&t.scan_tmp, bytes)) return BFGT_ERR_HIP;
    const TreeRules& r = t.rules;
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_bsz, t.boxsize.data(), sizeof(double) * (kTreeMaxLevel + 1),
                                  hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_xq, r.x.data(), sizeof(double) * t.n, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_U, r.U.data(), sizeof(double) * nn, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_P2C, r.P2C.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(
        hipMemcpyAsync(t.d_C2P, r.C2P.data(), sizeof(double) * r.C2P.size(), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_mask, r.mask.data(), sizeof(double) * nn, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_mask2, r.mask2.data(), sizeof(double) * nn, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_wq, r.wq.data(), sizeof(double) * nn, hipMemcpyHostToDevice, t.stream));
    return sync(t);
}

// exclusive sum of in[0..m) into out[0..m]; *total = sum
int scan(DeviceTree& t, const int* in, int* out, int m, int* total) {
    if (m == 0) {
        *total = 0;
        return 0;
    }
    size_t bytes = t.scan_bytes;
    BFGT_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t.scan_tmp, bytes, in, out, m, t.stream));
    int last[2];
    BFGT_HIP_CHECK(hipMemcpyAsync(&last[0], out + m - 1, sizeof(int), hipMemcpyDeviceToHost, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(&last[1], in + m - 1, sizeof(int), hipMemcpyDeviceToHost, t.stream));
    if (int rc = sync(t)) return rc;
    *total = last[0] + last[1];
    return 0;
}

// Sample fun on the boxes of a range into vals (per-box layout (nd, npbox)).
int sample(DeviceTree& t, Rng r, double* vals, bfgt_tree_fun_t fun, void* ctx) {
    if (r.m == 0) return 0;
    size_t npts = (size_t)r.m * t.npbox;
    if (npts > t.xy_cap) {
        if (t.d_xy) BFGT_HIP_CHECK(hipFree(t.d_xy));
        if (t.d_smp) BFGT_HIP_CHECK(hipFree(t.d_smp));
        t.d_xy = t.d_smp = nullptr;
        BFGT_HIP_CHECK(hipMalloc((void**)&t.d_xy, sizeof(double) * 2 * npts));
        BFGT_HIP_CHECK(hipMalloc((void**)&t.d_smp, sizeof(double) * t.nd * npts));
        t.xy_cap = npts;
    }
    k_box_points<<<blocks((long)npts, 256), 256, 0, t.stream>>>(r, t.n, t.A.lev, t.A.ctr, t.d_bsz, t.d_xq, t.d_xy);
    BFGT_HIP_CHECK(hipGetLastError());
    // the points are complete when the callback runs, whatever stream it uses
    if (int rc = sync(t)) return rc;
    if (r.list == nullptr) {
        // a contiguous range is contiguous in vals too: sample in place
        int rc = fun(ctx, t.nd, (int)npts, t.d_xy, vals + (size_t)r.first * t.npbox * t.nd, (void*)t.stream);
        if (rc) return set_error(BFGT_ERR_CALLBACK, "the user function callback returned " + std::to_string(rc));
        return 0;
    }
    int rc = fun(ctx, t.nd, (int)npts, t.d_xy, t.d_smp, (void*)t.stream);
    if (rc) return set_error(BFGT_ERR_CALLBACK, "the user function callback returned " + std::to_string(rc));
    int per = t.npbox * t.nd;
    k_scatter_vals<<<blocks((long)r.m * per, 256), 256, 0, t.stream>>>(r, per, t.d_smp, vals);
    BFGT_HIP_CHECK(hipGetLastError());
    return 0;
}

size_t err_lds(const DeviceTree& t) { return sizeof(double) * (5 * t.n * t.n + kTB); }

int flag_unresolved(DeviceTree& t, Rng r, const double* vals, const double* mask, double rsum, int iptype,
                    double eta, int big_split) {
    if (r.m == 0) return 0;
    k_unresolved<<<r.m, kTB, err_lds(t), t.stream>>>(r, t.n, t.nd, vals, t.d_U, mask, rsum, iptype, eta, t.A.lev,
                                                     t.d_bsz, t.d_thr, big_split, t.flag);
    BFGT_HIP_CHECK(hipGetLastError());
    return 0;
}

// Split the flagged leaves of the given ranges; children numbered from
// nboxes in the ranges' order.  Returns the number of split boxes in *nsplit
// (their parents in t.parents[0..nsplit)).
int split_ranges(DeviceTree& t, const std::vector<Rng>& rs, int* nsplit) {
    int total = 0;
    std::vector<int> tot(rs.size());
    // count first so that an overflow leaves the tree untouched
    for (size_t k = 0; k < rs.size(); k++) {
        if (rs[k].m == 0) {
            tot[k] = 0;
            continue;
        }
        k_take<<<blocks(rs[k].m, 256), 256, 0, t.stream>>>(rs[k], t.flag, t.A.nch, t.take);
        BFGT_HIP_CHECK(hipGetLastError());
        if (int rc = scan(t, t.take, t.off, rs[k].m, &tot[k])) return rc;
        if (tot[k] && (long)t.nboxes + 4L * (total + tot[k]) > t.cap)
            return set_error(BFGT_ERR_RESOURCE, "tree capacity of " + std::to_string(t.cap) + " boxes exceeded");
        if (tot[k]) {
            k_split<<<blocks(rs[k].m, 256), 256, 0, t.stream>>>(rs[k], t.take, t.off, t.nboxes + 4 * total, total, t.A,
                                                                t.d_bsz, t.parents);
            BFGT_HIP_CHECK(hipGetLastError());
        }
        total += tot[k];
    }
    *nsplit = total;
    return 0;
}

int interpolate_to_children(DeviceTree& t, int nparents, double* vals) {
    if (nparents == 0) return 0;
    size_t lds = sizeof(double) * 6 * t.n * t.n;
    k_to_children<<<nparents, kTB, lds, t.stream>>>(nparents, t.parents, t.n, t.nd, t.A.ch, vals, t.d_U, t.d_P2C);
    BFGT_HIP_CHECK(hipGetLastError());
    return 0;
}

// Renumber breadth first from the root; drops boxes no longer reachable.
// Moves vals (and fv when with_fv).
int renumber(DeviceTree& t, bool with_fv) {
    k_fill<int><<<blocks(t.nboxes, 256), 256, 0, t.stream>>>(t.newid, t.nboxes, -1);
    BFGT_HIP_CHECK(hipGetLastError());
    int zero = 0;
    BFGT_HIP_CHECK(hipMemcpyAsync(t.lista, &zero, sizeof(int), hipMemcpyHostToDevice, t.stream));
    int *cur = t.lista, *next = t.listb;
    int m = 1, base = 0, l = 0;
    t.lfirst.clear();
    t.lcount.clear();
    while (m > 0) {
        if (l > kTreeMaxLevel) return set_error(BFGT_ERR_TREE, "tree deeper than the level limit");
        k_level_ids<<<blocks(m, 256), 256, 0, t.stream>>>(m, cur, base, t.newid, t.A.nch, t.take);
        BFGT_HIP_CHECK(hipGetLastError());
        int tot = 0;
        if (int rc = scan(t, t.take, t.off, m, &tot)) return rc;
        if (tot) {
            k_level_next<<<blocks(m, 256), 256, 0, t.stream>>>(m, cur, t.off, t.A.ch, next);
            BFGT_HIP_CHECK(hipGetLastError());
        }
        t.lfirst.push_back(base);
        t.lcount.push_back(m);
        base += m;
        std::swap(cur, next);
        m = tot;
        l++;
    }
    int nold = t.nboxes;
    int per = t.npbox * t.nd;
    k_gather<<<blocks((long)nold * kTB, 256), 256, 0, t.stream>>>(nold, t.newid, per, t.A, t.B,
                                                                 with_fv ? t.fv : nullptr, t.fvB);
    BFGT_HIP_CHECK(hipGetLastError());
    std::swap(t.A, t.B);
    if (with_fv) std::swap(t.fv, t.fvB);
    t.nboxes = base;
    t.nlevels = (int)t.lcount.size() - 1;
    return sync(t);
}

int colleagues(DeviceTree& t) {
    int one = 1, root[9] = {0, -1, -1, -1, -1, -1, -1, -1, -1};
    BFGT_HIP_CHECK(hipMemcpyAsync(t.ncoll, &one, sizeof(int), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.coll, root, sizeof(root), hipMemcpyHostToDevice, t.stream));
    for (int l = 1; l <= t.nlevels; l++) {
        k_colleagues<<<blocks(t.lcount[l], 256), 256, 0, t.stream>>>(t.lfirst[l], t.lcount[l], l, t.iperiod, t.A.par,
                                                                     t.A.ch, t.A.cx, t.A.cy, t.ncoll, t.coll);
        BFGT_HIP_CHECK(hipGetLastError());
    }
    return sync(t);
}

// Level restriction of a renumbered tree.  New boxes get values from fun
// (when fun is set) or by interpolation from their parents.
int level_restrict(DeviceTree& t, bfgt_tree_fun_t fun, void* ctx) {
    if (t.nlevels < 2) return 0;
    long ks = (long)t.key_mask + 1;
    k_fill<unsigned long long><<<blocks(ks, 256), 256, 0, t.stream>>>(t.keys, ks, kEmptyKey);
    BFGT_HIP_CHECK(hipGetLastError());
    const long key_limit = ks / 2;
    long nkeys = 0;
    // (1) from the finest refined level up: cells of level k -> parents of
    // their neighbourhoods at level k-1
    int2 *cur = t.cells_a, *nxt = t.cells_b;
    int ncur = 0;  // cells of level k found while processing level k+1
    for (int k = t.nlevels - 1; k >= 1; k--) {
        // add this level's refined boxes (their keys go into the set here)
        BFGT_HIP_CHECK(hipMemcpyAsync(t.counter, &ncur, sizeof(int), hipMemcpyHostToDevice, t.stream));
        Rng r{t.lcount[k], t.lfirst[k], nullptr};
        k_refined_cells<<<blocks(r.m, 256), 256, 0, t.stream>>>(r, t.A.nch, t.A.lev, t.A.cx, t.A.cy, t.keys,
                                                                t.key_mask, cur, t.counter);
        BFGT_HIP_CHECK(hipGetLastError());
        int m = 0;
        if (int rc = read_int(t, t.counter, &m)) return rc;
        nkeys += m;
        if ((long)m * 4 > t.cap || nkeys > key_limit / 2)
            return set_error(BFGT_ERR_RESOURCE, "level restriction needs more than the tree capacity of " +
                                                    std::to_string(t.cap) + " boxes");
        int z = 0;
        BFGT_HIP_CHECK(hipMemcpyAsync(t.counter, &z, sizeof(int), hipMemcpyHostToDevice, t.stream));
        if (m) {
            k_close<<<blocks(m, 256), 256, 0, t.stream>>>(m, cur, k, t.iperiod, t.keys, t.key_mask, nxt, t.counter,
                                                          t.cap);
            BFGT_HIP_CHECK(hipGetLastError());
        }
        if (int rc = read_int(t, t.counter, &ncur)) return rc;
        nkeys += ncur;
        if ((long)ncur * 4 > t.cap || nkeys > key_limit / 2)
            return set_error(BFGT_ERR_RESOURCE, "level restriction needs more than the tree capacity of " +
                                                    std::to_string(t.cap) + " boxes");
        std::swap(cur, nxt);
    }
    // (2) top-down: split every leaf whose cell is in the set
    int new_first = -1, new_count = 0;  // boxes added at the current level
    for (int j = 0; j + 1 <= t.nlevels - 1; j++) {
        std::vector<Rng> rs;
        rs.push_back(Rng{t.lcount[j], t.lfirst[j], nullptr});
        if (new_count) rs.push_back(Rng{new_count, new_first, nullptr});
        for (const Rng& r : rs) {
            k_in_set<<<blocks(r.m, 256), 256, 0, t.stream>>>(r, t.A.nch, t.A.lev, t.A.cx, t.A.cy, t.keys, t.key_mask,
                                                             t.flag);
            BFGT_HIP_CHECK(hipGetLastError());
        }
        int ns = 0;
        int first_child = t.nboxes;
        if (int rc = split_ranges(t, rs, &ns)) return rc;
        t.nboxes += 4 * ns;
        new_first = first_child;
        new_count = 4 * ns;
        if (ns) {
            if (fun) {
                if (int rc = sample(t, Rng{new_count, new_first, nullptr}, t.A.vals, fun, ctx)) return rc;
            } else {
                if (int rc = interpolate_to_children(t, ns, t.A.vals)) return rc;
            }
        }
    }
    return renumber(t, false);
}

int set_thresholds(DeviceTree& t, const std::vector<double>& thr) {
    // the host copy stays alive until the next call (the copy is asynchronous)
    t.thr_host.assign(kTreeMaxLevel + 1, 0.0);
    for (size_t l = 0; l < t.thr_host.size() && l < thr.size(); l++) t.thr_host[l] = thr[l];
    BFGT_HIP_CHECK(hipMemcpyAsync(t.d_thr, t.thr_host.data(), sizeof(double) * t.thr_host.size(),
                                  hipMemcpyHostToDevice, t.stream));
    return 0;
}

double rsc0(int iptype, double boxlen) {
    if (iptype == 2) return std::sqrt(1.0 / (boxlen * boxlen));
    if (iptype == 1) return 1.0 / (boxlen * boxlen);
    return 1.0;
}

// Sum of per-box integrals of a range in the given order (host).
int range_integrals(DeviceTree& t, Rng r, double rsc, std::vector<double>& out) {
    out.assign(r.m, 0.0);
    if (r.m == 0) return 0;
    k_box_integral<<<blocks(r.m, 256), 256, 0, t.stream>>>(r, t.npbox, t.nd, t.A.vals, t.d_wq, rsc, t.iptype, t.d_red);
    BFGT_HIP_CHECK(hipGetLastError());
    BFGT_HIP_CHECK(hipMemcpyAsync(out.data(), t.d_red, sizeof(double) * r.m, hipMemcpyDeviceToHost, t.stream));
    return sync(t);
}

int tree_build(DeviceTree& t, double eps, bfgt_tree_fun_t fun, void* ctx, const double* rintl_in, int max_levels,
               double* rintl_out) {
    // root
    int zi = 0, mi = -1, m4[4] = {-1, -1, -1, -1};
    double z2[2] = {0.0, 0.0};
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.lev, &zi, sizeof(int), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.par, &mi, sizeof(int), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.nch, &zi, sizeof(int), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.ch, m4, sizeof(m4), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.cx, &zi, sizeof(int), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.cy, &zi, sizeof(int), hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.ctr, z2, sizeof(z2), hipMemcpyHostToDevice, t.stream));
    t.nboxes = 1;
    if (int rc = sample(t, Rng{1, 0, nullptr}, t.A.vals, fun, ctx)) return rc;
    const bool on_the_fly = rintl_in == nullptr;
    const double rs = rsc0(t.iptype, t.boxlen);
    std::vector<double> rintl(max_levels + 1, 0.0), rintbs_prev;
    double rint = 0.0;
    if (on_the_fly) {
        // vol_tree_mem2: integral over the root with rsc = boxlen^2 / 4
        if (int rc = range_integrals(t, Rng{1, 0, nullptr}, t.boxlen * t.boxlen / 4, rintbs_prev)) return rc;
        rint = t.iptype == 2 ? std::sqrt(rintbs_prev[0]) : rintbs_prev[0];
        rintl[0] = rint;
    } else {
        for (int l = 0; l <= max_levels; l++) rintl[l] = rintl_in[l];
    }
    int lfirst = 0, lcount = 1, lev = 0;
    for (; lev < max_levels; lev++) {
        std::vector<double> thr(kTreeMaxLevel + 1, 0.0);
        thr[lev] = eps * (rs * (on_the_fly ? rint : rintl[lev]));
        if (int rc = set_thresholds(t, thr)) return rc;
        Rng r{lcount, lfirst, nullptr};
        if (int rc = flag_unresolved(t, r, t.A.vals, t.d_mask, t.rules.rsum, t.iptype, t.eta, 1)) return rc;
        int ns = 0;
        int first_child = t.nboxes;
        if (int rc = split_ranges(t, {r}, &ns)) return rc;
        if (ns == 0) break;
        t.nboxes += 4 * ns;
        Rng kids{4 * ns, first_child, nullptr};
        if (int rc = sample(t, kids, t.A.vals, fun, ctx)) return rc;
        if (on_the_fly) {
            // update_rints: drop the split boxes' integrals, add the children's
            std::vector<int> par(ns);
            BFGT_HIP_CHECK(hipMemcpyAsync(par.data(), t.parents, sizeof(int) * ns, hipMemcpyDeviceToHost, t.stream));
            std::vector<double> kid;
            const double bsc = t.boxsize[lev + 1];
            if (int rc = range_integrals(t, kids, bsc * bsc / 4, kid)) return rc;
            if (t.iptype == 0) {
                for (double v : kid) rint = std::max(rint, v);
            } else {
                double acc = t.iptype == 2 ? rint * rint : rint;
                for (int i = 0; i < ns; i++) acc -= rintbs_prev[par[i] - lfirst];
                if (acc <= 0) acc = 0;
                for (double v : kid) acc += v;
                rint = t.iptype == 2 ? std::sqrt(acc) : acc;
            }
            rintl[lev + 1] = rint;
            rintbs_prev = kid;
        }
        lfirst = first_child;
        lcount = 4 * ns;
    }
    if (on_the_fly)
        for (int l = lev + 1; l <= max_levels; l++) rintl[l] = rint;
    if (rintl_out)
        for (int l = 0; l <= max_levels; l++) rintl_out[l] = rintl[l];
    if (int rc = renumber(t, false)) return rc;
    if (int rc = level_restrict(t, fun, ctx)) return rc;
    return colleagues(t);
}

// A tree given in the reference layout (1-based links, boxes in any order
// with the root first), with values vals(nd,npbox,nboxes) or zeros; checked
// for consistent links, levels and cells, then renumbered.
int tree_import(DeviceTree& t, int nboxes, int nlevels, const int* itree, const int* iptr, const double* centers,
                const double* boxsize, const double* vals) {
    if (nboxes < 1 || nboxes > t.cap) return set_error(BFGT_ERR_ARG, "nboxes must be in 1..max_boxes");
    if (nlevels < 0 || nlevels > kTreeMaxLevel) return set_error(BFGT_ERR_TREE, "nlevels outside 0..28");
    for (int l = 0; l <= nlevels; l++)
        if (boxsize[l] != t.boxsize[l]) return set_error(BFGT_ERR_TREE, "box sizes are not boxlen / 2^level");
    const int* ilevel = itree + iptr[1] - 1;
    const int* iparent = itree + iptr[2] - 1;
    const int* nchild = itree + iptr[3] - 1;
    const int* ichild = itree + iptr[4] - 1;
    std::vector<int> lev(nboxes), par(nboxes), nch(nboxes), ch(4 * (size_t)nboxes), cx(nboxes), cy(nboxes);
    std::vector<double> ctr(2 * (size_t)nboxes);
    for (int b = 0; b < nboxes; b++) {
        const int l = ilevel[b];
        if (l < 0 || l > nlevels) return set_error(BFGT_ERR_TREE, "box level outside 0..nlevels");
        const double bs = t.boxsize[l];
        const double fx = (centers[2 * b] + t.boxlen / 2) / bs - 0.5, fy = (centers[2 * b + 1] + t.boxlen / 2) / bs - 0.5;
        const long ix = std::lround(fx), iy = std::lround(fy);
        if (ix < 0 || iy < 0 || ix >= (1L << l) || iy >= (1L << l) || std::fabs(fx - ix) > 1e-6 ||
            std::fabs(fy - iy) > 1e-6)
            return set_error(BFGT_ERR_TREE, "box centre is not a cell centre of its level");
        lev[b] = l;
        cx[b] = (int)ix;
        cy[b] = (int)iy;
        ctr[2 * b] = centers[2 * b];
        ctr[2 * b + 1] = centers[2 * b + 1];
        par[b] = iparent[b] > 0 ? iparent[b] - 1 : -1;
        nch[b] = nchild[b];
        if (nch[b] != 0 && nch[b] != 4) return set_error(BFGT_ERR_TREE, "nchild must be 0 or 4");
        for (int j = 0; j < 4; j++) ch[4 * b + j] = nch[b] ? ichild[4 * (size_t)b + j] - 1 : -1;
    }
    if (lev[0] != 0 || par[0] != -1) return set_error(BFGT_ERR_TREE, "box 1 must be the root");
    for (int b = 0; b < nboxes; b++)
        for (int j = 0; nch[b] && j < 4; j++) {
            const int c = ch[4 * b + j];
            if (c < 0 || c >= nboxes || par[c] != b || lev[c] != lev[b] + 1 || cx[c] != 2 * cx[b] + (j & 1) ||
                cy[c] != 2 * cy[b] + (j >> 1))
                return set_error(BFGT_ERR_TREE, "inconsistent child links");
        }
    const size_t n = (size_t)nboxes;
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.lev, lev.data(), sizeof(int) * n, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.par, par.data(), sizeof(int) * n, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.nch, nch.data(), sizeof(int) * n, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.ch, ch.data(), sizeof(int) * 4 * n, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.cx, cx.data(), sizeof(int) * n, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.cy, cy.data(), sizeof(int) * n, hipMemcpyHostToDevice, t.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(t.A.ctr, ctr.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, t.stream));
    const size_t per = (size_t)t.npbox * t.nd;
    if (vals)
        BFGT_HIP_CHECK(hipMemcpyAsync(t.A.vals, vals, sizeof(double) * per * n, hipMemcpyHostToDevice, t.stream));
    else
        BFGT_HIP_CHECK(hipMemsetAsync(t.A.vals, 0, sizeof(double) * per * n, t.stream));
    t.nboxes = nboxes;
    // the host vectors above must outlive the copies
    if (int rc = sync(t)) return rc;
    if (int rc = renumber(t, false)) return rc;
    if (t.nboxes != nboxes) return set_error(BFGT_ERR_TREE, "boxes not reachable from the root");
    return colleagues(t);
}

// Leaves of the current (renumbered) tree into t.lista; returns the count.
int leaf_list(DeviceTree& t, int* count) {
    k_leaf_flags<<<blocks(t.nboxes, 256), 256, 0, t.stream>>>(t.nboxes, t.A.nch, t.take);
    BFGT_HIP_CHECK(hipGetLastError());
    if (int rc = scan(t, t.take, t.off, t.nboxes, count)) return rc;
    k_compact<<<blocks(t.nboxes, 256), 256, 0, t.stream>>>(t.nboxes, t.take, t.off, t.lista);
    BFGT_HIP_CHECK(hipGetLastError());
    return 0;
}

int tree_adapt(DeviceTree& t, double eps, const double* rintl, bfgt_tree_fun_t fun, void* ctx, int max_levels) {
    if (!t.fv) {
        if (dalloc(t, &t.fv, (size_t)t.cap * t.npbox * t.nd) || dalloc(t, &t.fvB, (size_t)t.cap * t.npbox * t.nd))
            return BFGT_ERR_HIP;
    }
    const double reps = eps * rsc0(t.iptype, t.boxlen);
    std::vector<double> thr(kTreeMaxLevel + 1, 0.0);
    for (int l = 0; l <= std::min(max_levels, kTreeMaxLevel); l++) thr[l] = reps * rintl[l];
    if (int rc = set_thresholds(t, thr)) return rc;
    // fun on the leaves; split flags of the leaves
    int nleaf = 0;
    if (int rc = leaf_list(t, &nleaf)) return rc;
    k_fill<int><<<blocks(t.nboxes, 256), 256, 0, t.stream>>>(t.flag, t.nboxes, 0);
    BFGT_HIP_CHECK(hipGetLastError());
    Rng leaves{nleaf, 0, t.lista};
    if (int rc = sample(t, leaves, t.fv, fun, ctx)) return rc;
    if (int rc = flag_unresolved(t, leaves, t.fv, t.d_mask, t.rules.rsum, t.iptype, 1.0, 0)) return rc;
    // split level by level, down to max_levels
    int new_first = 0, new_count = 0;
    const int nl0 = t.nlevels;
    for (int l = 0;; l++) {
        std::vector<Rng> rs;
        if (l <= nl0) rs.push_back(Rng{t.lcount[l], t.lfirst[l], nullptr});
        if (new_count) rs.push_back(Rng{new_count, new_first, nullptr});
        if (rs.empty()) break;
        int ns = 0, first_child = t.nboxes;
        if (int rc = split_ranges(t, rs, &ns)) return rc;
        if (ns && l + 1 > max_levels)
            return set_error(BFGT_ERR_RESOURCE, "re-meshing needs more than " + std::to_string(max_levels) + " levels");
        t.nboxes += 4 * ns;
        new_first = first_child;
        new_count = 4 * ns;
        if (ns) {
            if (int rc = interpolate_to_children(t, ns, t.A.vals)) return rc;
            Rng kids{new_count, new_first, nullptr};
            if (int rc = sample(t, kids, t.fv, fun, ctx)) return rc;
            if (int rc = flag_unresolved(t, kids, t.fv, t.d_mask, t.rules.rsum, t.iptype, 1.0, 0)) return rc;
        }
        if (l >= nl0 && ns == 0) break;
    }
    if (int rc = renumber(t, false)) return rc;
    // fun on every leaf of the refined tree (vol_tree_set_vals), then merging
    if (int rc = leaf_list(t, &nleaf)) return rc;
    if (int rc = sample(t, Rng{nleaf, 0, t.lista}, t.fv, fun, ctx)) return rc;
    size_t lds = sizeof(double) * (5 * t.n * t.n + (t.n / 2) * t.n * 2 + kTB);
    for (int l = t.nlevels - 1; l >= 0; l--) {
        Rng r{t.lcount[l], t.lfirst[l], nullptr};
        k_merge_children<<<r.m, kTB, lds, t.stream>>>(r, t.n, t.nd, t.A, t.fv, t.d_U, t.d_C2P, t.d_mask2,
                                                      t.rules.rsum2, reps);
        BFGT_HIP_CHECK(hipGetLastError());
    }
    if (int rc = renumber(t, false)) return rc;
    if (int rc = level_restrict(t, nullptr, nullptr)) return rc;
    return colleagues(t);
}

}  // namespace

struct TreeParams {
    int nd, ipoly, iperiod, norder, iptype;
    double eta, boxlen;
    int max_boxes;
    void* stream;
};

// A new tree object set up from the parameters and filled by `fill`; the
// handle is returned only when filling succeeded.
template <typename Fill>
int tree_new(bfgt_tree_t* out, const TreeParams& tp, Fill fill) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(BFGT_ERR_NODEVICE, "no GPU visible");
    std::unique_ptr<bfgt_tree> h;
    int rc;
    try {
        h.reset(new bfgt_tree());
        DeviceTree& t = h->t;
        BFGT_HIP_CHECK(hipGetDevice(&t.device));
        if (tp.stream) {
            t.stream = (hipStream_t)tp.stream;
        } else {
            BFGT_HIP_CHECK(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking));
            t.own_stream = true;
        }
        t.nd = tp.nd;
        t.n = tp.norder;
        t.npbox = tp.norder * tp.norder;
        t.ipoly = tp.ipoly;
        t.iperiod = tp.iperiod ? 1 : 0;
        t.iptype = tp.iptype;
        t.eta = tp.eta;
        t.boxlen = tp.boxlen;
        t.cap = tp.max_boxes;
        t.boxsize.assign(kTreeMaxLevel + 1, 0.0);
        t.boxsize[0] = tp.boxlen;
        for (int l = 1; l <= kTreeMaxLevel; l++) t.boxsize[l] = t.boxsize[l - 1] / 2;
        t.rules = make_rules(tp.norder, tp.ipoly, tp.iptype);
        rc = fill(t);
    } catch (const std::exception& e) {
        rc = set_error(BFGT_ERR_RESOURCE, std::string("tree set-up: ") + e.what());
    }
    if (rc) return rc;
    *out = h.release();
    return BFGT_OK;
}

}  // namespace bfgt

// --------------------------------------------------------------------------
// C ABI

using bfgt::set_error;

extern "C" {

int bfgt_tree_build(bfgt_tree_t* out, int nd, int ipoly, int iperiod, double eps, double boxlen, int norder,
                    int iptype, double eta, bfgt_tree_fun_t fun, void* ctx, const double* rintl_in, int max_levels,
                    int max_boxes, double* rintl_out, void* stream) {
    if (!out) return set_error(BFGT_ERR_ARG, "null tree pointer");
    *out = nullptr;
    if (!fun) return set_error(BFGT_ERR_ARG, "null user function");
    if (nd < 1) return set_error(BFGT_ERR_ARG, "nd must be >= 1");
    if (ipoly != 0 && ipoly != 1) return set_error(BFGT_ERR_IPOLY, "ipoly must be 0 or 1");
    if (norder < 2 || norder > bfgt::kMaxOrder) return set_error(BFGT_ERR_NORDER, "norder must be in 2..32");
    if (iptype < 0 || iptype > 2) return set_error(BFGT_ERR_ARG, "iptype must be 0, 1 or 2");
    if (!(eps > 0) || !(boxlen > 0)) return set_error(BFGT_ERR_ARG, "eps and boxlen must be positive");
    if (max_levels < 0 || max_levels > bfgt::kTreeMaxLevel)
        return set_error(BFGT_ERR_ARG, "max_levels must be in 0..28");
    if (max_boxes < 1) return set_error(BFGT_ERR_ARG, "max_boxes must be positive");
    bfgt::TreeParams tp{nd, ipoly, iperiod, norder, iptype, eta, boxlen, max_boxes, stream};
    return bfgt::tree_new(out, tp, [&](bfgt::DeviceTree& t) {
        if (int rc = bfgt::tree_alloc(t)) return rc;
        return bfgt::tree_build(t, eps, fun, ctx, rintl_in, max_levels, rintl_out);
    });
}

int bfgt_tree_import(bfgt_tree_t* out, int nd, int ipoly, int iperiod, int norder, int iptype, double eta,
                     int nboxes, int nlevels, const int* itree, const int* iptr, const double* centers,
                     const double* boxsize, const double* vals, int max_boxes, void* stream) {
    if (!out) return set_error(BFGT_ERR_ARG, "null tree pointer");
    *out = nullptr;
    if (!itree || !iptr || !centers || !boxsize) return set_error(BFGT_ERR_ARG, "null tree array");
    if (nd < 1) return set_error(BFGT_ERR_ARG, "nd must be >= 1");
    if (ipoly != 0 && ipoly != 1) return set_error(BFGT_ERR_IPOLY, "ipoly must be 0 or 1");
    if (norder < 2 || norder > bfgt::kMaxOrder) return set_error(BFGT_ERR_NORDER, "norder must be in 2..32");
    if (iptype < 0 || iptype > 2) return set_error(BFGT_ERR_ARG, "iptype must be 0, 1 or 2");
    if (!(boxsize[0] > 0)) return set_error(BFGT_ERR_TREE, "boxsize(0) must be positive");
    if (max_boxes < nboxes) return set_error(BFGT_ERR_ARG, "max_boxes below nboxes");
    bfgt::TreeParams tp{nd, ipoly, iperiod, norder, iptype, eta, boxsize[0], max_boxes, stream};
    return bfgt::tree_new(out, tp, [&](bfgt::DeviceTree& t) {
        if (int rc = bfgt::tree_alloc(t)) return rc;
        return bfgt::tree_import(t, nboxes, nlevels, itree, iptr, centers, boxsize, vals);
    });
}

int bfgt_tree_adapt(bfgt_tree_t tree, double eps, const double* rintl, bfgt_tree_fun_t fun, void* ctx,
                    int max_levels) {
    if (!tree || !fun || !rintl) return set_error(BFGT_ERR_ARG, "null tree, function or rintl");
    if (!(eps > 0)) return set_error(BFGT_ERR_ARG, "eps must be positive");
    if (max_levels < 0 || max_levels > bfgt::kTreeMaxLevel)
        return set_error(BFGT_ERR_ARG, "max_levels must be in 0..28");
    if (tree->t.n % 2) return set_error(BFGT_ERR_NORDER, "re-meshing needs an even norder");
    try {
        return bfgt::tree_adapt(tree->t, eps, rintl, fun, ctx, max_levels);
    } catch (const std::exception& e) {
        return set_error(BFGT_ERR_RESOURCE, std::string("re-meshing: ") + e.what());
    }
}

int bfgt_tree_query(bfgt_tree_t tree, int* iv, double** d_vals) {
    if (!tree) return set_error(BFGT_ERR_ARG, "null tree");
    const bfgt::DeviceTree& t = tree->t;
    if (iv) {
        int nleaf = 0;
        std::vector<int> nch(t.nboxes);
        if (t.nboxes) BFGT_HIP_CHECK(hipMemcpy(nch.data(), t.A.nch, sizeof(int) * t.nboxes, hipMemcpyDeviceToHost));
        for (int v : nch) nleaf += v == 0;
        int vals[10] = {t.nboxes, t.nlevels, 2 * (t.nlevels + 1) + 17 * t.nboxes, nleaf, t.cap, t.nd, t.n,
                        t.ipoly,  t.iperiod, t.iptype};
        std::memcpy(iv, vals, sizeof(vals));
    }
    if (d_vals) *d_vals = t.A.vals;
    return BFGT_OK;
}

int bfgt_tree_export(bfgt_tree_t tree, int* itree, int* iptr, double* centers, double* boxsize, double* vals) {
    if (!tree || !itree || !iptr || !centers || !boxsize) return set_error(BFGT_ERR_ARG, "null argument");
    bfgt::DeviceTree& t = tree->t;
    const int nb = t.nboxes, nl = t.nlevels;
    int* d = nullptr;
    BFGT_HIP_CHECK(hipMalloc((void**)&d, sizeof(int) * 17 * (size_t)nb));
    double* dc = nullptr;
    if (hipMalloc((void**)&dc, sizeof(double) * 2 * (size_t)nb) != hipSuccess) {
        (void)hipFree(d);
        return set_error(BFGT_ERR_HIP, "hipMalloc in bfgt_tree_export");
    }
    bfgt::k_export<<<bfgt::blocks(nb, 256), 256, 0, t.stream>>>(nb, t.A, t.ncoll, t.coll, d, d + nb, d + 2 * nb,
                                                                d + 3 * nb, d + 7 * nb, d + 8 * nb, dc);
    hipError_t e = hipGetLastError();
    int lad = 2 * (nl + 1);
    if (e == hipSuccess) e = hipMemcpyAsync(itree + lad, d, sizeof(int) * 17 * (size_t)nb, hipMemcpyDeviceToHost, t.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(centers, dc, sizeof(double) * 2 * (size_t)nb, hipMemcpyDeviceToHost, t.stream);
    if (e == hipSuccess && vals)
        e = hipMemcpyAsync(vals, t.A.vals, sizeof(double) * (size_t)nb * t.npbox * t.nd, hipMemcpyDeviceToHost,
                           t.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t.stream);
    (void)hipFree(d);
    (void)hipFree(dc);
    if (e != hipSuccess) return set_error(BFGT_ERR_HIP, std::string("bfgt_tree_export: ") + hipGetErrorString(e));
    for (int l = 0; l <= nl; l++) {
        itree[2 * l] = t.lfirst[l] + 1;
        itree[2 * l + 1] = t.lfirst[l] + t.lcount[l];
        boxsize[l] = t.boxsize[l];
    }
    iptr[0] = 1;
    iptr[1] = lad + 1;
    iptr[2] = iptr[1] + nb;
    iptr[3] = iptr[2] + nb;
    iptr[4] = iptr[3] + nb;
    iptr[5] = iptr[4] + 4 * nb;
    iptr[6] = iptr[5] + nb;
    iptr[7] = iptr[6] + 9 * nb;
    return BFGT_OK;
}

int bfgt_tree_plan_create(bfgt_plan_t* out, bfgt_tree_t tree, double delta, double eps, int ifpgh, int precision,
                          void* stream) {
    if (!out || !tree) return set_error(BFGT_ERR_ARG, "null argument");
    const bfgt::DeviceTree& t = tree->t;
    std::vector<int> itree, iptr(8);
    std::vector<double> centers, boxsize;
    try {
        itree.assign(2 * (t.nlevels + 1) + 17 * (size_t)t.nboxes, 0);
        centers.assign(2 * (size_t)t.nboxes, 0.0);
        boxsize.assign(t.nlevels + 1, 0.0);
    } catch (const std::exception& e) {
        return set_error(BFGT_ERR_RESOURCE, std::string("tree export: ") + e.what());
    }
    if (int rc = bfgt_tree_export(tree, itree.data(), iptr.data(), centers.data(), boxsize.data(), nullptr)) return rc;
    return bfgt_plan_create(out, t.nd, 2, delta, eps, t.ipoly, t.iperiod, t.n, t.nboxes, t.nlevels, (int)itree.size(),
                            itree.data(), iptr.data(), centers.data(), boxsize.data(), ifpgh, precision, stream);
}

int bfgt_tree_destroy(bfgt_tree_t tree) {
    if (!tree) return BFGT_OK;
    (void)hipStreamSynchronize(tree->t.stream);
    delete tree;
    return BFGT_OK;
}

}  // extern "C"
