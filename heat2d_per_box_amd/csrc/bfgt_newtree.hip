// Tree refinement inside the box FGT (ifnewtree = 1, bfgtmain STEPS 8-10,
// bfgt/boxfgt.f:1172-1483): after the transform the tree is refined where
// pot is not resolved, coarsened where it is over-resolved, and made 2:1
// level-restricted again, with pot/grad/hess carried to the new boxes.
//
// The numerics run on the GPU, on the value arrays the transform has just
// written: tensor-product coefficients and tail norms of every leaf, the
// potential of new boxes evaluated (not interpolated) from the cutoff
// boxes' local expansions and the near field of the original leaf's list 1
// with tables for targets refined up to three times, coefficient
// interpolation child -> parent and parent -> child, and the renumberings.
// The bookkeeping -- which boxes split or merge, in what order they are
// numbered, the colleague lists -- runs on the host over the caller's
// itree, following the reference's rules, since the numbering of the output
// tree is part of the contract.  Refinement is done one generation at a
// time (children of the boxes split in one pass are evaluated together);
// the reference's single pass over appended boxes numbers them the same way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "bfgt_internal.h"

namespace bfgt {
namespace {

constexpr int kNT = 64;  // threads of the per-box kernels

// coefficients of channel f (of nf per point) of one box's values vb:
// c[kx + n ky] = sum U(kx, ix) U(ky, iy) v(ix, iy); v, t LDS scratch of n*n.
// Block-wide; ends synchronised.
__device__ void coefs_of(int n, int nf, int f, const double* vb, const double* U, double* v, double* t, double* c) {
    const int nn = n * n;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) v[q] = vb[(long)q * nf + f];
    __syncthreads();
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        const int kx = q % n, iy = q / n;
        double s = 0.0;
        for (int ix = 0; ix < n; ix++) s += U[kx * n + ix] * v[ix + n * iy];
        t[q] = s;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        const int kx = q % n, ky = q / n;
        double s = 0.0;
        for (int iy = 0; iy < n; iy++) s += U[ky * n + iy] * t[kx + n * iy];
        c[q] = s;
    }
    __syncthreads();
}

// block-wide sum through red[blockDim.x]; ends synchronised
__device__ double block_sum(double a, double* red) {
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// out[i] = sum over the points and densities of box[i] of pot^2 w (the
// product quadrature weights): the box's share of the l2 norm before the
// level's area factor (treedata_lpnorm, common/tree_data_routs_nd.f:1131)
__global__ __launch_bounds__(kNT) void k_leaf_sq(const int* __restrict__ boxes, int n, int nd,
                                                 const double* __restrict__ pot, const double* __restrict__ wq,
                                                 double* __restrict__ out) {
    __shared__ double red[kNT];
    const int b = boxes[blockIdx.x];
    const int nn = n * n;
    double a = 0.0;
    for (int e = threadIdx.x; e < nn * nd; e += blockDim.x) {
        const double v = pot[(long)b * nn * nd + e];
        a += v * v * wq[e / nd];
    }
    const double s = block_sum(a, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Coefficients of the listed boxes' values (nf functions per point), stored
// in coefs when given, and err[i] = max_f |masked coefficients|_2 * rscale[i]
// / rsum when err is given (fun_err with iptype 2,
// common/tree_vol_coeffs_nd.f:855-935).
__global__ __launch_bounds__(kNT) void k_coef_err(const int* __restrict__ boxes, const double* __restrict__ rscale,
                                                  int n, int nf, const double* __restrict__ vals,
                                                  const double* __restrict__ Ug, const double* __restrict__ maskg,
                                                  double rsum, double* __restrict__ coefs, double* __restrict__ err) {
    extern __shared__ double sh[];
    const int nn = n * n;
    double *U = sh, *mask = U + nn, *v = mask + nn, *t = v + nn, *c = t + nn, *red = c + nn;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) {
        U[q] = Ug[q];
        mask[q] = maskg[q];
    }
    __syncthreads();
    const int b = boxes[blockIdx.x];
    double e = 0.0;
    for (int f = 0; f < nf; f++) {
        coefs_of(n, nf, f, vals + (long)b * nn * nf, U, v, t, c);
        double a = 0.0;
        for (int q = threadIdx.x; q < nn; q += blockDim.x) {
            if (coefs) coefs[((long)b * nn + q) * nf + f] = c[q];
            a += mask[q] * c[q] * c[q];
        }
        e = fmax(e, sqrt(block_sum(a, red)));
    }
    if (threadIdx.x == 0 && err) err[blockIdx.x] = e * rscale[blockIdx.x] / rsum;
}

// One family per entry: fam[5 i] the parent, fam[5 i + 1 + j] its children
// in the reference's order (j: bit 0 = x plus side, bit 1 = y plus side).

// Children -> parent: each child's expansion (coefficients from src, or
// from its values when src holds values) at the parent's nodes in the
// child's quadrant (C2P [s][i][k], i < n/2), written into the parent's
// values dst (vol_tree_coarsen, common/tree_vol_coeffs_nd.f:2282-2340).
__global__ __launch_bounds__(kNT) void k_c2p(const int* __restrict__ fam, int n, int nf, const double* __restrict__ src,
                                             int src_coefs, const double* __restrict__ Ug,
                                             const double* __restrict__ C2Pg, double* __restrict__ dst) {
    extern __shared__ double sh[];
    const int nn = n * n, n2 = n / 2;
    double *U = sh, *P = U + nn, *v = P + 2 * n2 * n, *t = v + nn, *c = t + nn;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) U[q] = Ug[q];
    for (int q = threadIdx.x; q < 2 * n2 * n; q += blockDim.x) P[q] = C2Pg[q];
    __syncthreads();
    const int b = fam[5 * blockIdx.x];
    for (int f = 0; f < nf; f++) {
        for (int j = 0; j < 4; j++) {
            const int cj = fam[5 * blockIdx.x + 1 + j];
            if (src_coefs) {
                for (int q = threadIdx.x; q < nn; q += blockDim.x) c[q] = src[((long)cj * nn + q) * nf + f];
                __syncthreads();
            } else {
                coefs_of(n, nf, f, src + (long)cj * nn * nf, U, v, t, c);
            }
            const double* Px = P + (j & 1) * n2 * n;
            const double* Py = P + (j >> 1) * n2 * n;
            for (int q = threadIdx.x; q < n2 * n; q += blockDim.x) {
                const int i = q % n2, ky = q / n2;
                double s = 0.0;
                for (int kx = 0; kx < n; kx++) s += Px[i * n + kx] * c[kx + n * ky];
                t[q] = s;
            }
            __syncthreads();
            const int ox = (j & 1) * n2, oy = (j >> 1) * n2;
            for (int q = threadIdx.x; q < n2 * n2; q += blockDim.x) {
                const int i = q % n2, i2 = q / n2;
                double s = 0.0;
                for (int ky = 0; ky < n; ky++) s += Py[i2 * n + ky] * t[i + n2 * ky];
                dst[((long)b * nn + (ox + i) + n * (oy + i2)) * nf + f] = s;
            }
            __syncthreads();
        }
    }
}

// Parent -> children: the parent's expansion (coefficients from src, or
// from its values) at each child's nodes (P2C [s][i][k]), written into the
// children's values dst (vol_tree_refine_boxes_flag_interp,
// common/tree_vol_coeffs_nd.f:3170-3205).
__global__ __launch_bounds__(kNT) void k_p2c(const int* __restrict__ fam, int n, int nf, const double* __restrict__ src,
                                             int src_coefs, const double* __restrict__ Ug,
                                             const double* __restrict__ P2Cg, double* __restrict__ dst) {
    extern __shared__ double sh[];
    const int nn = n * n;
    double *U = sh, *P = U + nn, *v = P + 2 * nn, *t = v + nn, *c = t + nn;
    for (int q = threadIdx.x; q < nn; q += blockDim.x) U[q] = Ug[q];
    for (int q = threadIdx.x; q < 2 * nn; q += blockDim.x) P[q] = P2Cg[q];
    __syncthreads();
    const int b = fam[5 * blockIdx.x];
    for (int f = 0; f < nf; f++) {
        if (src_coefs) {
            for (int q = threadIdx.x; q < nn; q += blockDim.x) c[q] = src[((long)b * nn + q) * nf + f];
            __syncthreads();
        } else {
            coefs_of(n, nf, f, src + (long)b * nn * nf, U, v, t, c);
        }
        for (int j = 0; j < 4; j++) {
            const int cj = fam[5 * blockIdx.x + 1 + j];
            const double* Px = P + (j & 1) * nn;
            const double* Py = P + (j >> 1) * nn;
            for (int q = threadIdx.x; q < nn; q += blockDim.x) {
                const int ix = q % n, ky = q / n;
                double s = 0.0;
                for (int kx = 0; kx < n; kx++) s += Px[ix * n + kx] * c[kx + n * ky];
                t[q] = s;
            }
            __syncthreads();
            for (int q = threadIdx.x; q < nn; q += blockDim.x) {
                const int ix = q % n, iy = q / n;
                double s = 0.0;
                for (int ky = 0; ky < n; ky++) s += Py[iy * n + ky] * t[ix + n * ky];
                dst[((long)cj * nn + q) * nf + f] = s;
            }
            __syncthreads();
        }
    }
}

// dst[p] = src[order[p]], per_box doubles per box
__global__ void k_gather(long cnt, const int* __restrict__ order, int per_box, const double* __restrict__ src,
                         double* __restrict__ dst) {
    const long tot = cnt * per_box;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const long p = e / per_box;
        dst[e] = src[(long)order[p] * per_box + e % per_box];
    }
}

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// device buffers of one refining call
struct NtDev {
    double *coefs = nullptr, *tmp = nullptr, *dbl = nullptr, *err = nullptr;
    double *U = nullptr, *P2C = nullptr, *C2P = nullptr, *mask = nullptr, *wq = nullptr;
    int* ibuf = nullptr;
    double* tloc = nullptr;
    int* indloc = nullptr;
    cplx* loc = nullptr;
    PwLeaf* pw = nullptr;
    int *ntgt = nullptr, *nlev = nullptr, *nptr = nullptr;
    NearPair* npair = nullptr;
    size_t npair_cap = 0;
    std::vector<void*> all;
    ~NtDev() {
        for (void* q : all) (void)hipFree(q);
    }
    template <typename T>
    int get(T** out, size_t count) {
        if (hipMalloc((void**)out, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
            return set_error(BFGT_ERR_HIP, "hipMalloc failed for the refinement buffers");
        all.push_back(*out);
        return BFGT_OK;
    }
};

struct Refiner {
    Plan& p;
    NewTreeIO& io;
    NtDev d;
    TreeRules R;
    TreeView T;
    int nbmax = 0;  // boxes the caller's arrays hold
    int nlev = 0;   // current number of levels
    int n = 0, nn = 0, nd = 0, nf = 0, iperiod = 0;
    double reps = 0, rsum = 1;
    int half = 0;  // offsets of the refined near-field tables: -half..half
    int anc0 = 0;  // first box of the cutoff level (cutoff box a is box anc0 + a)

    Refiner(Plan& p_, NewTreeIO& io_) : p(p_), io(io_) {}

    void view() {
        const int* ip = io.iptr;
        int* it = io.itree;
        T.nlevels = nlev;
        T.laddr = it + ip[0] - 1;
        T.ilevel = it + ip[1] - 1;
        T.iparent = it + ip[2] - 1;
        T.nchild = it + ip[3] - 1;
        T.ichild = it + ip[4] - 1;
        T.ncoll = it + ip[5] - 1;
        T.coll = it + ip[6] - 1;
        T.centers = io.centers;
        T.boxsize = io.boxsize;
        T.nboxes = T.laddr[2 * nlev + 1];
    }
    int level(int b) const { return T.ilevel[b]; }
    int count() const { return T.laddr[2 * nlev + 1]; }

    int sync() {
        BFGT_HIP_CHECK(hipStreamSynchronize(p.stream));
        return BFGT_OK;
    }
    // every box id a kernel receives is checked against the arrays first
    int check_ids(const std::vector<int>& v) const {
        for (int b : v)
            if (b < 0 || b >= nbmax) return set_error(BFGT_ERR_TREE, "refinement: box id out of range");
        return BFGT_OK;
    }
    int put_ints(const std::vector<int>& v) {
        int rc = check_ids(v);
        if (rc) return rc;
        BFGT_HIP_CHECK(hipMemcpyAsync(d.ibuf, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, p.stream));
        return BFGT_OK;
    }

    int setup();
    int coef_err(const std::vector<int>& boxes, const std::vector<double>& rscale, std::vector<double>* err);
    int permute(const std::vector<int>& order, std::vector<int>* iflag);
    int step8();
    int evaluate_new(const std::vector<int>& made, const std::vector<int>& orig, const std::vector<char>& pwexp,
                     const std::vector<std::vector<int>>& list1, const std::vector<int>& anc);
    int step9();
    int refine_flagged(int first, int last, int ilev, int* nboxes, std::vector<int>& iflag);
    void updateflags(int first, int last, int curlev, std::vector<int>& iflag);
    int reorg_tail(const std::vector<int>& tail, std::vector<int>& iflag);
    int colleagues();
    int step10();
};

int Refiner::setup() {
    n = p.n;
    nn = n * n;
    nd = p.nd;
    nf = nd;
    iperiod = p.iperiod_eff;
    nlev = *io.nlevels;
    // room in the caller's arrays: iptr sizes every box array
    nbmax = io.nbmax;
    const int widths[6] = {1, 1, 1, 4, 1, 9};
    for (int k = 1; k <= 6; k++) {
        const long end = k < 6 ? io.iptr[k + 1] : (long)io.ltree + 1;
        nbmax = (int)std::min<long>(nbmax, (end - io.iptr[k]) / widths[k - 1]);
    }
    view();
    R = make_rules(n, p.ipoly, 2);
    rsum = R.rsum;
    const size_t cap = (size_t)nbmax;
    int rc;
    if ((rc = d.get(&d.coefs, cap * nn * nd))) return rc;
    if ((rc = d.get(&d.tmp, cap * nn * nd * std::max(1, p.ifpgh)))) return rc;
    if ((rc = d.get(&d.dbl, cap))) return rc;
    if ((rc = d.get(&d.err, cap))) return rc;
    if ((rc = d.get(&d.ibuf, 5 * cap + 8))) return rc;
    if ((rc = d.get(&d.U, R.U.size()))) return rc;
    if ((rc = d.get(&d.P2C, R.P2C.size()))) return rc;
    if ((rc = d.get(&d.C2P, R.C2P.size()))) return rc;
    if ((rc = d.get(&d.mask, R.mask.size()))) return rc;
    if ((rc = d.get(&d.wq, R.wq.size()))) return rc;
    BFGT_HIP_CHECK(hipMemcpyAsync(d.U, R.U.data(), sizeof(double) * R.U.size(), hipMemcpyHostToDevice, p.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(d.P2C, R.P2C.data(), sizeof(double) * R.P2C.size(), hipMemcpyHostToDevice, p.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(d.C2P, R.C2P.data(), sizeof(double) * R.C2P.size(), hipMemcpyHostToDevice, p.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(d.mask, R.mask.data(), sizeof(double) * R.mask.size(), hipMemcpyHostToDevice, p.stream));
    BFGT_HIP_CHECK(hipMemcpyAsync(d.wq, R.wq.data(), sizeof(double) * R.wq.size(), hipMemcpyHostToDevice, p.stream));
    return BFGT_OK;
}

int Refiner::coef_err(const std::vector<int>& boxes, const std::vector<double>& rscale, std::vector<double>* err) {
    const int m = (int)boxes.size();
    if (m == 0) return BFGT_OK;
    int rc = put_ints(boxes);
    if (rc) return rc;
    BFGT_HIP_CHECK(hipMemcpyAsync(d.dbl, rscale.data(), sizeof(double) * m, hipMemcpyHostToDevice, p.stream));
    const size_t sh = sizeof(double) * (5 * nn + kNT);
    k_coef_err<<<m, kNT, sh, p.stream>>>(d.ibuf, d.dbl, n, nd, io.pot, d.U, d.mask, rsum, d.coefs,
                                          err ? d.err : nullptr);
    BFGT_HIP_CHECK(hipGetLastError());
    if (err) {
        err->resize(m);
        BFGT_HIP_CHECK(hipMemcpyAsync(err->data(), d.err, sizeof(double) * m, hipMemcpyDeviceToHost, p.stream));
    }
    return sync();
}

// Renumber: new position q takes box order[q]; parents and children are
// mapped through the inverse (-1 kept); pot/grad/hess and the coefficients
// move on the device.  laddr is the caller's business.
int Refiner::permute(const std::vector<int>& order, std::vector<int>* iflag) {
    const int cnt = (int)order.size();
    std::vector<int> inv(nbmax, -1);
    for (int q = 0; q < cnt; q++) inv[order[q]] = q;
    std::vector<int> il(T.ilevel, T.ilevel + nbmax), ip(T.iparent, T.iparent + nbmax), nc(T.nchild, T.nchild + nbmax),
        ic(T.ichild, T.ichild + 4 * (size_t)nbmax);
    std::vector<double> c2(io.centers, io.centers + 2 * (size_t)nbmax);
    int rc = check_ids(order);
    if (rc) return rc;
    std::vector<int> fl;
    if (iflag) fl = *iflag;
    for (int q = 0; q < cnt; q++) {
        const int o = order[q];
        T.ilevel[q] = il[o];
        T.nchild[q] = nc[o];
        T.iparent[q] = ip[o] > 0 ? inv[ip[o] - 1] + 1 : -1;
        for (int j = 0; j < 4; j++) T.ichild[4 * q + j] = ic[4 * o + j] > 0 ? inv[ic[4 * o + j] - 1] + 1 : -1;
        io.centers[2 * q] = c2[2 * o];
        io.centers[2 * q + 1] = c2[2 * o + 1];
        if (iflag) (*iflag)[q] = fl[o];
    }
    BFGT_HIP_CHECK(hipMemcpyAsync(d.ibuf, order.data(), sizeof(int) * cnt, hipMemcpyHostToDevice, p.stream));
    double* arr[4] = {io.pot, d.coefs, p.ifpgh >= 2 ? io.grad : nullptr, p.ifpgh >= 3 ? io.hess : nullptr};
    const int per[4] = {nn * nd, nn * nd, 2 * nn * nd, 3 * nn * nd};
    for (int a = 0; a < 4; a++) {
        if (!arr[a]) continue;
        const long tot = (long)cnt * per[a];
        const int blocks = (int)std::min<long>((tot + 255) / 256, 8192);
        k_gather<<<std::max(blocks, 1), 256, 0, p.stream>>>(cnt, d.ibuf, per[a], arr[a], d.tmp);
        BFGT_HIP_CHECK(hipGetLastError());
        BFGT_HIP_CHECK(hipMemcpyAsync(arr[a], d.tmp, sizeof(double) * tot, hipMemcpyDeviceToDevice, p.stream));
    }
    return sync();
}

// pot/grad/hess of the boxes just created: the cutoff box's local expansion
// evaluated on them (the reference shifts the parent's local expansion to
// the child, gnd_shiftpw_loc, and evaluates it, gnd_pw2pgh -- the same sum),
// plus the near field of the original leaf's list 1 with the child's level
// tables (bfgt/boxfgt.f:1324-1360)
int Refiner::evaluate_new(const std::vector<int>& made, const std::vector<int>& orig, const std::vector<char>& pwexp,
                          const std::vector<std::vector<int>>& list1, const std::vector<int>& anc) {
    std::vector<PwLeaf> pw;
    std::vector<int> tgt, tlev, ptr(1, 0);
    std::vector<NearPair> pairs;
    const double sq = std::sqrt(p.delta0), bs0 = io.boxsize[0];
        const int nsrc = p.nboxes;
    for (int jb : made) {
        const int o = orig[jb];
        if (jb < 0 || jb >= nbmax || o < 0 || o >= nsrc)
            return set_error(BFGT_ERR_TREE, "refinement: box id out of range");
        const int jl = level(jb);
        if (pwexp[o] == 1 && p.has_pw) {
            const int a = anc[o];
            const int ab = anc0 + a;
            PwLeaf lf{};
            lf.box = jb;
            lf.level = jl;
            lf.anc = a;
            lf.dx = (io.centers[2 * jb] - io.centers[2 * ab]) / sq;
            lf.dy = (io.centers[2 * jb + 1] - io.centers[2 * ab + 1]) / sq;
            pw.push_back(lf);
        }
        if (!list1[o].empty()) {
            for (int kb : list1[o]) {
                if (kb < 0 || kb >= nsrc) return set_error(BFGT_ERR_TREE, "refinement: source id out of range");
                // gnd_find_loctab_ind with mrefinelev = 3 (bfgt/bfgt_local.f:1729-1765)
                const double u = io.boxsize[level(kb)] / 32.0;
                int ix[2];
                for (int k = 0; k < 2; k++) {
                    double dx = (io.centers[2 * jb + k] - io.centers[2 * kb + k]) / u;
                    if (iperiod == 1) {
                        const double dp = dx - bs0 / u, dm = dx + bs0 / u;
                        if (std::fabs(dx) > std::fabs(dp)) dx = dp;
                        if (std::fabs(dx) > std::fabs(dm)) dx = dm;
                    }
                    ix[k] = (int)dx;
                }
                if (std::abs(ix[0]) > half || std::abs(ix[1]) > half)
                    return set_error(BFGT_ERR_RESOURCE, "refined box outside the range of the near-field tables");
                pairs.push_back(NearPair{kb, ix[0], ix[1], 0});
            }
            tgt.push_back(jb);
            tlev.push_back(jl);
            ptr.push_back((int)pairs.size());
        }
    }
    if (!pw.empty()) {
        BFGT_HIP_CHECK(hipMemcpyAsync(d.pw, pw.data(), sizeof(PwLeaf) * pw.size(), hipMemcpyHostToDevice, p.stream));
        int rc = launch_pw_eval_list(p, d.pw, (int)pw.size(), d.loc, io.pot, io.grad, io.hess);
        if (rc) return rc;
    }
    if (!tgt.empty()) {
        if (pairs.size() > d.npair_cap) {
            int rc = d.get(&d.npair, pairs.size());
            if (rc) return rc;
            d.npair_cap = pairs.size();
        }
        BFGT_HIP_CHECK(hipMemcpyAsync(d.ntgt, tgt.data(), sizeof(int) * tgt.size(), hipMemcpyHostToDevice, p.stream));
        BFGT_HIP_CHECK(hipMemcpyAsync(d.nlev, tlev.data(), sizeof(int) * tlev.size(), hipMemcpyHostToDevice, p.stream));
        BFGT_HIP_CHECK(hipMemcpyAsync(d.nptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, p.stream));
        BFGT_HIP_CHECK(
            hipMemcpyAsync(d.npair, pairs.data(), sizeof(NearPair) * pairs.size(), hipMemcpyHostToDevice, p.stream));
        int rc = launch_near_list(p, (int)tgt.size(), d.ntgt, d.nlev, d.nptr, d.npair, d.tloc, half, io.fvals,
                                  io.pot, io.grad, io.hess);
        if (rc) return rc;
    }
    return sync();
}

// STEP 8 (bfgt/boxfgt.f:1172-1412): leaves above the finest level whose
// pot is not resolved to eps * (l2 norm of pot) are split; their children
// are evaluated and split in turn, at most imaxrefinelev generations below
// the original leaf, and a generation g only when 2 g <= maxrefinelev (the
// reference's irefinelev counts two per generation).  Deviation: a box at
// the caller's finest level is not split -- the reference would write
// boxes at a level that boxsize, laddr and its tables do not hold.
int Refiner::step8() {
    const int nboxes0 = count();
    // the l2 norm of pot over the leaves (treedata_lpnorm)
    std::vector<int> leaves;
    std::vector<double> area, rs;
    for (int l = 0; l <= nlev; l++)
        for (int b = T.first(l); b <= T.last(l); b++)
            if (T.nchild[b] == 0) {
                leaves.push_back(b);
                area.push_back((io.boxsize[l] / 2) * (io.boxsize[l] / 2));
                rs.push_back(io.boxsize[l] / 2);
            }
    const int m = (int)leaves.size();
    std::vector<double> sq(m);
    int rc = put_ints(leaves);
    if (rc) return rc;
    k_leaf_sq<<<m, kNT, 0, p.stream>>>(d.ibuf, n, nd, io.pot, d.wq, d.err);
    BFGT_HIP_CHECK(hipGetLastError());
    BFGT_HIP_CHECK(hipMemcpyAsync(sq.data(), d.err, sizeof(double) * m, hipMemcpyDeviceToHost, p.stream));
    rc = sync();
    if (rc) return rc;
    double rnorm = 0;
    for (int i = 0; i < m; i++) rnorm += sq[i] * area[i];
    rnorm = std::sqrt(rnorm);
    const double bs0 = io.boxsize[0];
    reps = p.eps * std::sqrt(1.0 / (bs0 * bs0)) * rnorm;
    // coefficients of every leaf; errors of those above the finest level
    std::vector<double> err;
    if ((rc = coef_err(leaves, rs, &err))) return rc;
    std::vector<char> flag(nbmax, 0);
    std::vector<int> imax(nbmax, 0), irl(nbmax, 0), orig(nbmax, -1), gen(nbmax, 0);
    int maxrl = 0;
    for (int i = 0; i < m; i++) {
        const int b = leaves[i];
        if (level(b) >= nlev) continue;
        if (err[i] > reps) {
            flag[b] = 1;
            imax[b] = (int)std::ceil(std::log(err[i] / reps) / std::log(2.0) / n);
        }
        maxrl = std::max(maxrl, imax[b]);
    }
    // lists of the transform's tree: PW leaves, modified list 1, cutoff boxes
    std::vector<char> pwexp = pwexp_flags(T, p.npwlevel);
    std::vector<std::vector<int>> list1 = modified_list1(T, p.npwlevel, pwexp, iperiod);
    std::vector<int> anc(nboxes0, -1);
    if (p.has_pw) {
        anc0 = T.first(p.nlevstart);
        for (int b = 0; b < nboxes0; b++) {
            if (level(b) < p.nlevstart) continue;
            int a = b;
            while (level(a) > p.nlevstart) a = T.iparent[a] - 1;
            anc[b] = a - anc0;
        }
    }
    std::vector<int> cur;
    for (int b = 0; b < nboxes0; b++)
        if (flag[b]) cur.push_back(b);
    int nnew = 0, ncap = 0;
    while (!cur.empty()) {
        std::vector<int> made;
        for (int ib : cur) {
            if (!(flag[ib] && irl[ib] <= maxrl)) continue;
            const int il = level(ib);
            if (il >= nlev) {
                ncap++;
                continue;
            }
            irl[ib]++;
            const double h = io.boxsize[il] / 4;
            T.nchild[ib] = 4;
            for (int j = 0; j < 4; j++) {
                const int jb = nboxes0 + nnew++;
                if (jb >= nbmax) return set_error(BFGT_ERR_RESOURCE, "the refined tree outgrows nboxes");
                irl[jb] = irl[ib] + 1;
                orig[jb] = ib < nboxes0 ? ib : orig[ib];
                gen[jb] = gen[ib] + 1;
                io.centers[2 * jb] = io.centers[2 * ib] + ((j & 1) ? h : -h);
                io.centers[2 * jb + 1] = io.centers[2 * ib + 1] + ((j >> 1) ? h : -h);
                T.iparent[jb] = ib + 1;
                T.nchild[jb] = 0;
                for (int l = 0; l < 4; l++) T.ichild[4 * jb + l] = -1;
                T.ichild[4 * ib + j] = jb + 1;
                T.ilevel[jb] = il + 1;
                made.push_back(jb);
            }
        }
        if (made.empty()) break;
        // orig[] of an original box is itself
        for (int b : cur)
            if (b < nboxes0) orig[b] = b;
        if ((rc = evaluate_new(made, orig, pwexp, list1, anc))) return rc;
        std::vector<double> rsm(made.size());
        for (size_t i = 0; i < made.size(); i++) rsm[i] = io.boxsize[level(made[i]) - 1] / 4;
        if ((rc = coef_err(made, rsm, &err))) return rc;
        std::vector<int> next;
        for (size_t i = 0; i < made.size(); i++) {
            const int jb = made[i];
            if (err[i] > reps && gen[jb] <= maxrl && gen[jb] <= imax[orig[jb]]) {
                flag[jb] = 1;
                next.push_back(jb);
            }
        }
        cur.swap(next);
    }
    io.nnew = nnew;
    io.ncapped = ncap;
    if (nnew == 0) return BFGT_OK;
    // fgt_vol_tree_reorg (common/tree_vol_coeffs_nd.f:1973-2157): per level
    // the old boxes, then the new ones in creation order
    const int cnt = nboxes0 + nnew;
    std::vector<int> order;
    order.reserve(cnt);
    order.push_back(0);
    std::vector<int> first(nlev + 1), last(nlev + 1);
    for (int l = 1; l <= nlev; l++) {
        first[l] = (int)order.size() + 1;
        for (int b = T.first(l); b <= T.last(l); b++) order.push_back(b);
        for (int b = nboxes0; b < cnt; b++)
            if (level(b) == l) order.push_back(b);
        last[l] = (int)order.size();
    }
    if ((rc = permute(order, nullptr))) return rc;
    for (int l = 1; l <= nlev; l++) {
        T.laddr[2 * l] = first[l];
        T.laddr[2 * l + 1] = last[l];
    }
    view();
    return BFGT_OK;
}

// STEP 9 (vol_tree_coarsen + fgt_vol_tree_reorg_after_coarsen,
// common/tree_vol_coeffs_nd.f:2158-2570), from the finest level up: a box
// whose four children are leaves gets their pot interpolated onto its
// nodes; when that is resolved to reps (no scale) the children are dropped
// and grad/hess are interpolated too.
int Refiner::step9() {
    int rc;
    std::vector<char> del(nbmax, 0);
    int ndel = 0;
    std::vector<int> fam, parents;
    std::vector<double> ones, err;
    const size_t shc = sizeof(double) * (4 * nn + 2 * (n / 2) * n);
    for (int il = nlev - 1; il >= 0; il--) {
        fam.clear();
        parents.clear();
        for (int b = T.first(il); b <= T.last(il); b++) {
            if (T.nchild[b] != 4) continue;
            bool leafy = true;
            for (int j = 0; j < 4; j++) leafy = leafy && T.nchild[T.ichild[4 * b + j] - 1] == 0;
            if (!leafy) continue;
            parents.push_back(b);
            fam.push_back(b);
            for (int j = 0; j < 4; j++) fam.push_back(T.ichild[4 * b + j] - 1);
        }
        if (parents.empty()) continue;
        const int m = (int)parents.size();
        if ((rc = put_ints(fam))) return rc;
        k_c2p<<<m, kNT, shc, p.stream>>>(d.ibuf, n, nd, d.coefs, 1, d.U, d.C2P, io.pot);
        BFGT_HIP_CHECK(hipGetLastError());
        ones.assign(m, 1.0);
        if ((rc = coef_err(parents, ones, &err))) return rc;
        fam.clear();
        std::vector<int> merged;
        for (int i = 0; i < m; i++)
            if (err[i] <= reps) {
                const int b = parents[i];
                merged.push_back(b);
                fam.push_back(b);
                for (int j = 0; j < 4; j++) fam.push_back(T.ichild[4 * b + j] - 1);
            }
        if (merged.empty()) continue;
        const int mm = (int)merged.size();
        if ((rc = put_ints(fam))) return rc;
        if (p.ifpgh >= 2) k_c2p<<<mm, kNT, shc, p.stream>>>(d.ibuf, n, 2 * nd, io.grad, 0, d.U, d.C2P, io.grad);
        if (p.ifpgh >= 3) k_c2p<<<mm, kNT, shc, p.stream>>>(d.ibuf, n, 3 * nd, io.hess, 0, d.U, d.C2P, io.hess);
        BFGT_HIP_CHECK(hipGetLastError());
        if ((rc = sync())) return rc;
        for (int b : merged) {
            for (int j = 0; j < 4; j++) {
                const int c = T.ichild[4 * b + j] - 1;
                T.iparent[c] = -1;
                T.nchild[c] = 0;
                for (int l = 0; l < 4; l++) T.ichild[4 * c + l] = -1;
                del[c] = 1;
                ndel++;
            }
            T.nchild[b] = 0;
            for (int j = 0; j < 4; j++) T.ichild[4 * b + j] = -1;
        }
    }
    if (ndel == 0) return BFGT_OK;
    std::vector<int> order;
    order.push_back(0);
    std::vector<int> first(nlev + 1), last(nlev + 1);
    for (int l = 1; l <= nlev; l++) {
        first[l] = (int)order.size() + 1;
        for (int b = T.first(l); b <= T.last(l); b++)
            if (!del[b]) order.push_back(b);
        last[l] = (int)order.size();
    }
    if ((rc = permute(order, nullptr))) return rc;
    for (int l = 1; l <= nlev; l++) {
        T.laddr[2 * l] = first[l];
        T.laddr[2 * l + 1] = last[l];
    }
    int nl = nlev;
    for (int l = nlev; l >= 0; l--)
        if (T.laddr[2 * l + 1] < T.laddr[2 * l]) nl--;
    nlev = nl;
    view();
    return BFGT_OK;
}

// vol_tree_refine_boxes_flag_interp (common/tree_vol_coeffs_nd.f:3127-3223):
// flagged boxes of [first, last] split in order, children numbered from
// *nboxes; pot from the parent's coefficients, grad/hess from theirs;
// children of flag-1 boxes become flag++ (3), of flag+ boxes 0
int Refiner::refine_flagged(int first, int last, int il, int* nboxes, std::vector<int>& iflag) {
    std::vector<int> fam, kids;
    const double h = io.boxsize[il + 1] / 2;
    for (int b = first; b <= last; b++) {
        if (iflag[b] <= 0) continue;
        if (*nboxes + 4 > nbmax) return set_error(BFGT_ERR_RESOURCE, "the rebalanced tree outgrows nboxes");
        T.nchild[b] = 4;
        fam.push_back(b);
        for (int j = 0; j < 4; j++) {
            const int c = (*nboxes)++;
            io.centers[2 * c] = io.centers[2 * b] + ((j & 1) ? h : -h);
            io.centers[2 * c + 1] = io.centers[2 * b + 1] + ((j >> 1) ? h : -h);
            T.iparent[c] = b + 1;
            T.nchild[c] = 0;
            for (int l = 0; l < 4; l++) T.ichild[4 * c + l] = -1;
            T.ichild[4 * b + j] = c + 1;
            T.ilevel[c] = il + 1;
            if (iflag[b] == 1) iflag[c] = 3;
            if (iflag[b] == 2) iflag[c] = 0;
            fam.push_back(c);
            kids.push_back(c);
        }
    }
    if (fam.empty()) return BFGT_OK;
    const int m = (int)fam.size() / 5;
    int rc;
    if ((rc = put_ints(fam))) return rc;
    const size_t shp = sizeof(double) * 6 * nn;  // U, P2C (2), v, t, c
    k_p2c<<<m, kNT, shp, p.stream>>>(d.ibuf, n, nd, d.coefs, 1, d.U, d.P2C, io.pot);
    if (p.ifpgh >= 2) k_p2c<<<m, kNT, shp, p.stream>>>(d.ibuf, n, 2 * nd, io.grad, 0, d.U, d.P2C, io.grad);
    if (p.ifpgh >= 3) k_p2c<<<m, kNT, shp, p.stream>>>(d.ibuf, n, 3 * nd, io.hess, 0, d.U, d.P2C, io.hess);
    BFGT_HIP_CHECK(hipGetLastError());
    if ((rc = sync())) return rc;  // the family list is reused by coef_err's upload
    return coef_err(kids, std::vector<double>(kids.size(), 1.0), nullptr);
}

// vol_updateflags (common/tree_vol_coeffs_nd.f:1733-1850) on [first, last]
void Refiner::updateflags(int first, int last, int cl, std::vector<int>& iflag) {
    const double distest = 1.05 * (io.boxsize[cl] + io.boxsize[cl + 1]) / 2.0, bs0 = io.boxsize[0];
    for (int b = first; b <= last; b++) {
        if (iflag[b] != 3) continue;
        iflag[b] = 0;
        for (int i = 0; i < T.ncoll[b] && iflag[b] == 0; i++) {
            const int jb = T.coll[9 * b + i] - 1;
            for (int j = 0; j < 4; j++) {
                int kb = T.ichild[4 * jb + j];
                if (kb <= 0) continue;
                kb -= 1;
                if (T.nchild[kb] <= 0) continue;
                int ict = 0;
                for (int k = 0; k < 2; k++)
                    if (per_dist(io.centers[2 * kb + k] - io.centers[2 * b + k], iperiod, bs0) <= distest) ict++;
                if (ict == 2) {
                    iflag[b] = 1;
                    break;
                }
            }
        }
    }
}

// vol_tree_reorg_pgh (common/tree_vol_coeffs_nd.f:2916-3126): levels 0 and
// 1 stay; every deeper level lists its laddr range, then its tail
int Refiner::reorg_tail(const std::vector<int>& tail, std::vector<int>& iflag) {
    std::vector<int> order;
    for (int l = 0; l <= std::min(nlev, 1); l++)
        for (int b = T.first(l); b <= T.last(l); b++) order.push_back(b);
    std::vector<int> first(nlev + 1), last(nlev + 1);
    for (int l = 2; l <= nlev; l++) {
        first[l] = (int)order.size() + 1;
        for (int b = T.first(l); b <= T.last(l); b++) order.push_back(b);
        for (int b = tail[2 * l] - 1; b <= tail[2 * l + 1] - 1; b++) order.push_back(b);
        last[l] = (int)order.size();
    }
    int rc = permute(order, &iflag);
    if (rc) return rc;
    for (int l = 2; l <= nlev; l++) {
        T.laddr[2 * l] = first[l];
        T.laddr[2 * l + 1] = last[l];
    }
    view();
    return BFGT_OK;
}

int Refiner::colleagues() {
    view();
    computecoll(T, iperiod);
    return BFGT_OK;
}

// STEP 10 (bfgt/boxfgt.f:1459-1483): colleagues, then
// vol_tree_fix_lr_interp (common/tree_vol_coeffs_nd.f:2571-2915)
int Refiner::step10() {
    colleagues();
    const double bs0 = io.boxsize[0];
    std::vector<int> iflag(nbmax, 0);
    int nboxes = count();
    auto touch = [&](int a, int b, double dt) {
        for (int k = 0; k < 2; k++)
            if (!(per_dist(io.centers[2 * a + k] - io.centers[2 * b + k], iperiod, bs0) <= dt)) return false;
        return true;
    };
    // flag: leaves touching a box two levels finer (seen from its grandparent)
    for (int il = nlev; il >= 2; il--) {
        const double dt = 1.05 * (io.boxsize[il - 1] + io.boxsize[il - 2]) / 2.0;
        for (int b = T.first(il); b <= T.last(il); b++) {
            const int dad = T.iparent[b] - 1, gd = T.iparent[dad] - 1;
            for (int i = 0; i < T.ncoll[gd]; i++) {
                const int jb = T.coll[9 * gd + i] - 1;
                if (T.nchild[jb] == 0 && iflag[jb] == 0 && touch(jb, dad, dt)) iflag[jb] = 1;
            }
        }
    }
    // flag+: leaves bigger than and touching a flagged or flag+ box
    for (int il = nlev; il >= 1; il--) {
        const double dt = 1.05 * (io.boxsize[il] + io.boxsize[il - 1]) / 2.0;
        for (int b = T.first(il); b <= T.last(il); b++) {
            if (iflag[b] != 1 && iflag[b] != 2) continue;
            const int dad = T.iparent[b] - 1;
            for (int i = 0; i < T.ncoll[dad]; i++) {
                const int jb = T.coll[9 * dad + i] - 1;
                if (T.nchild[jb] == 0 && iflag[jb] == 0 && touch(jb, b, dt)) iflag[jb] = 2;
            }
        }
    }
    std::vector<int> tail(2 * (nlev + 1));
    auto reset = [&]() {
        for (int l = 0; l <= nlev; l++) {
            tail[2 * l] = 0;
            tail[2 * l + 1] = -1;
        }
    };
    reset();
    int rc;
    for (int il = 1; il <= nlev - 2; il++) {
        tail[2 * (il + 1)] = nboxes + 1;
        if ((rc = refine_flagged(T.first(il), T.last(il), il, &nboxes, iflag))) return rc;
        tail[2 * (il + 1) + 1] = nboxes;
    }
    if ((rc = reorg_tail(tail, iflag))) return rc;
    colleagues();
    for (int i = 0; i < nboxes; i++)
        if (iflag[i] != 3) iflag[i] = 0;
    reset();
    // flag++ boxes: split while they still touch a box two levels finer,
    // level by level, the boxes created meanwhile included
    for (int il = 2; il <= nlev - 2; il++) {
        updateflags(T.first(il), T.last(il), il, iflag);
        updateflags(tail[2 * il] - 1, tail[2 * il + 1] - 1, il, iflag);
        tail[2 * (il + 1)] = nboxes + 1;
        if ((rc = refine_flagged(T.first(il), T.last(il), il, &nboxes, iflag))) return rc;
        if ((rc = refine_flagged(tail[2 * il] - 1, tail[2 * il + 1] - 1, il, &nboxes, iflag))) return rc;
        tail[2 * (il + 1) + 1] = nboxes;
        // colleagues of the boxes just created: children of the parent's
        // colleagues that touch
        for (int b = tail[2 * (il + 1)] - 1; b <= tail[2 * (il + 1) + 1] - 1; b++) {
            T.ncoll[b] = 0;
            const int dad = T.iparent[b] - 1;
            for (int i = 0; i < T.ncoll[dad]; i++) {
                const int jb = T.coll[9 * dad + i] - 1;
                for (int j = 0; j < 4; j++) {
                    const int kb = T.ichild[4 * jb + j];
                    if (kb > 0 && touch(kb - 1, b, 1.05 * io.boxsize[il + 1]) && T.ncoll[b] < 9)
                        T.coll[9 * b + T.ncoll[b]++] = kb;
                }
            }
        }
    }
    if ((rc = reorg_tail(tail, iflag))) return rc;
    colleagues();
    return BFGT_OK;
}

}  // namespace

int newtree_refine(Plan& p, double eps, NewTreeIO& io) {
    (void)eps;
    Refiner r(p, io);
    int rc = r.setup();
    if (rc) return rc;
    if (p.has_pw) {
        if ((rc = r.d.get(&r.d.loc, (size_t)p.nanc * p.nd * p.nexp))) return rc;
        if ((rc = unfold_local(p, r.d.loc))) return rc;
    }
    if ((rc = r.d.get(&r.d.tloc, (size_t)(p.nlevels + 1) * 193 * 3 * p.n * p.n))) return rc;
    if ((rc = r.d.get(&r.d.indloc, (size_t)(p.nlevels + 1) * 193 * 2 * (p.n + 1)))) return rc;
    if ((rc = build_loctab_refined(p, 3, r.d.tloc, r.d.indloc, &r.half))) return rc;
    if ((rc = r.d.get(&r.d.pw, (size_t)r.nbmax))) return rc;
    if ((rc = r.d.get(&r.d.ntgt, (size_t)r.nbmax))) return rc;
    if ((rc = r.d.get(&r.d.nlev, (size_t)r.nbmax))) return rc;
    if ((rc = r.d.get(&r.d.nptr, (size_t)r.nbmax + 1))) return rc;
    const int nboxes0 = r.count();
    auto t0 = Clock::now();
    if ((rc = r.step8())) return rc;
    auto t1 = Clock::now();
    io.t_refine = secs(t0, t1);
    if (io.nnew > 0) {
        if (!(io.nnew < 0.1 * nboxes0)) {
            if ((rc = r.step9())) return rc;
        }
        auto t2 = Clock::now();
        io.t_coarsen = secs(t1, t2);
        if ((rc = r.step10())) return rc;
        io.t_rebalance = secs(t2, Clock::now());
    }
    *io.nlevels = r.nlev;
    // the final tree's nodes, for extra targets
    const int nb = r.count();
    std::vector<TreeNode> nodes(nb);
    for (int b = 0; b < nb; b++) {
        TreeNode& t = nodes[b];
        t.cx = io.centers[2 * b];
        t.cy = io.centers[2 * b + 1];
        for (int j = 0; j < 4; j++) t.child[j] = r.T.nchild[b] ? r.T.ichild[4 * b + j] - 1 : -1;
        t.level = r.T.ilevel[b];
        t.pad = 0;
    }
    BFGT_HIP_CHECK(hipMalloc(&io.d_nodes, sizeof(TreeNode) * nb));
    BFGT_HIP_CHECK(hipMemcpyAsync(io.d_nodes, nodes.data(), sizeof(TreeNode) * nb, hipMemcpyHostToDevice, p.stream));
    BFGT_HIP_CHECK(hipStreamSynchronize(p.stream));
    return BFGT_OK;
}

}  // namespace bfgt
