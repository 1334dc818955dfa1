// Extra-target evaluation of the box FGT on the GPU (FP64).
//
// After the transform, the reference interpolates the leaf arrays at user
// targets (bfgtmain STEP 11 with iffast=0, bfgt/boxfgt.f:1494-1563): each
// channel's leaf values are turned into tensor-polynomial coefficients
// (treedata_trans_nd, common/tree_data_routs_nd.f:16-90), every target is
// located in its leaf by descending the tree with ties to the plus side
// (pts_tree_sort / find_childbox_ind, common/pts_tree_nd.f:650-741,
// 1226-1253), and the expansion is evaluated at the target scaled to
// [-1,1]^2 (treedata_evalt_nd :710-806).  Here one thread owns one target:
// it descends the tree, forms the 1-D weights
//   w(i) = sum_k U(k, i) P_k(s)        (U: values -> coefficients)
// for both axes from the recurrence of the node family, and contracts them
// with the leaf values directly -- the same interpolant without an
// intermediate coefficient array, so nothing is allocated per call.
#include <hip/hip_runtime.h>

#include "bfgt_internal.h"

namespace bfgt {

namespace {

constexpr int kTargetThreads = 256;

// sum over a leaf's n x n grid of wx(i) wy(j) v[(i + n j) * nf + f]
template <int NMAX>
__device__ inline double contract(const double* __restrict__ v, int n, int nf, int f, const double* wx,
                                  const double* wy) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NMAX; j++) {
        if (j < n) {
            double r = 0.0;
#pragma unroll
            for (int i = 0; i < NMAX; i++)
                if (i < n) r = fma(wx[i], v[(size_t)(i + n * j) * nf + f], r);
            s = fma(wy[j], r, s);
        }
    }
    return s;
}

template <int NMAX>
__global__ __launch_bounds__(kTargetThreads) void k_targets(
    int ntarg, const double* __restrict__ targs, int ifpghtarg, int n, int ipoly, int nd, int maxdepth,
    const TreeNode* __restrict__ nodes, const double* __restrict__ U, const double* __restrict__ boxsize,
    const double* __restrict__ pot, const double* __restrict__ grad, const double* __restrict__ hess,
    double* __restrict__ pote, double* __restrict__ grade, double* __restrict__ hesse) {
    __shared__ double sU[NMAX * NMAX];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) sU[e] = U[e];
    __syncthreads();
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= ntarg) return;  // no barrier after this point
    const double x = targs[2 * it], y = targs[2 * it + 1];
    int b = 0;
    TreeNode nb = nodes[0];
    for (int depth = 0; depth < maxdepth && nb.child[0] >= 0; depth++) {
        int k = (x - nb.cx >= 0.0 ? 1 : 0) + (y - nb.cy >= 0.0 ? 2 : 0);
        b = nb.child[k];
        nb = nodes[b];
    }
    const double sc = 2.0 / boxsize[nb.level];
    const double xs = (x - nb.cx) * sc, ys = (y - nb.cy) * sc;
    double wx[NMAX], wy[NMAX];
#pragma unroll
    for (int i = 0; i < NMAX; i++) wx[i] = wy[i] = 0.0;
    // P_k / T_k by the three-term recurrence (common/legeexps.f:597-634, common/chebexps.f:594-626)
    double px0 = 1.0, px1 = xs, py0 = 1.0, py1 = ys;
#pragma unroll
    for (int k = 0; k < NMAX; k++) {
        if (k < n) {
            double pk = k == 0 ? 1.0 : px1, qk = k == 0 ? 1.0 : py1;
            if (k >= 2) {
                double a, c;
                if (ipoly == 0) {
                    a = ((2 * k - 1) * xs * px1 - (k - 1) * px0) / k;
                    c = ((2 * k - 1) * ys * py1 - (k - 1) * py0) / k;
                } else {
                    a = 2 * xs * px1 - px0;
                    c = 2 * ys * py1 - py0;
                }
                px0 = px1;
                px1 = a;
                py0 = py1;
                py1 = c;
                pk = a;
                qk = c;
            }
#pragma unroll
            for (int i = 0; i < NMAX; i++) {
                if (i < n) {
                    wx[i] = fma(pk, sU[k * n + i], wx[i]);
                    wy[i] = fma(qk, sU[k * n + i], wy[i]);
                }
            }
        }
    }
    const size_t npbox = (size_t)n * n;
    const double* vp = pot + (size_t)b * npbox * nd;
    for (int f = 0; f < nd; f++) pote[(size_t)it * nd + f] = contract<NMAX>(vp, n, nd, f, wx, wy);
    if (ifpghtarg >= 2) {
        const double* vg = grad + (size_t)b * npbox * 2 * nd;
        for (int f = 0; f < 2 * nd; f++) grade[(size_t)it * 2 * nd + f] = contract<NMAX>(vg, n, 2 * nd, f, wx, wy);
    }
    if (ifpghtarg >= 3) {
        const double* vh = hess + (size_t)b * npbox * 3 * nd;
        for (int f = 0; f < 3 * nd; f++) hesse[(size_t)it * 3 * nd + f] = contract<NMAX>(vh, n, 3 * nd, f, wx, wy);
    }
}

template <int NMAX>
int launch_targets_n(Plan& p, int ntarg, const double* targs, int ifpghtarg, const double* pot, const double* grad,
                     const double* hess, double* pote, double* grade, double* hesse) {
    int blocks = (ntarg + kTargetThreads - 1) / kTargetThreads;
    k_targets<NMAX><<<blocks, kTargetThreads, 0, p.stream>>>(ntarg, targs, ifpghtarg, p.n, p.ipoly, p.nd,
                                                             p.nlevels + 1, p.d_tnodes, p.d_nodes + p.n,
                                                             p.d_boxsize, pot, grad, hess, pote, grade, hesse);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// Tensor-grid node coordinates of every leaf, point i + n j at
// (cx + x_i b/2, cy + x_j b/2) with x the node rule on [-1, 1]: the points
// at which the callers evaluate their forcing (hpots/hpots2dadap.f:296-306,
// xtarg = centers + xref * boxsize/2).  Non-leaf boxes get zeros.
__global__ void k_leaf_points(int nboxes, int n, const TreeNode* __restrict__ nodes, const double* __restrict__ xq,
                              const double* __restrict__ boxsize, double* __restrict__ xy) {
    const size_t tot = (size_t)nboxes * n * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(e / ((size_t)n * n));
        const int pt = (int)(e % ((size_t)n * n));
        const TreeNode nb = nodes[b];
        double x = 0.0, y = 0.0;
        if (nb.child[0] < 0) {
            const double h = 0.5 * boxsize[nb.level];
            x = nb.cx + xq[pt % n] * h;
            y = nb.cy + xq[pt / n] * h;
        }
        xy[2 * e] = x;
        xy[2 * e + 1] = y;
    }
}

}  // namespace

int launch_leaf_points(Plan& p, double* xy) {
    size_t tot = (size_t)p.nboxes * p.n * p.n;
    int blocks = (int)((tot + 255) / 256 < 4096 ? (tot + 255) / 256 : 4096);
    k_leaf_points<<<blocks, 256, 0, p.stream>>>(p.nboxes, p.n, p.d_tnodes, p.d_nodes, p.d_boxsize, xy);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_targets(Plan& p, int ntarg, const double* targs, int ifpghtarg, const double* pot, const double* grad,
                   const double* hess, double* pote, double* grade, double* hesse) {
    if (ntarg == 0) return BFGT_OK;
    if (p.n <= 8) return launch_targets_n<8>(p, ntarg, targs, ifpghtarg, pot, grad, hess, pote, grade, hesse);
    if (p.n <= 16) return launch_targets_n<16>(p, ntarg, targs, ifpghtarg, pot, grad, hess, pote, grade, hesse);
    return launch_targets_n<32>(p, ntarg, targs, ifpghtarg, pot, grad, hess, pote, grade, hesse);
}

}  // namespace bfgt
