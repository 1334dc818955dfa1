// Internal declarations of the MI355X box FGT: the plan object, device
// buffers, and the kernel launch wrappers shared by the translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/boxfgt_hip.h"

namespace bfgt {

constexpr int kMaxOrder = 32;  // largest leaf order n the kernels accept
constexpr int kMaxNpw = 200;   // plane-wave length limit (the reference's ws/ts(200))
constexpr int kNq = 50;        // Gauss-Legendre nodes of the near-field quadrature branch

struct cplx {
    double x, y;
};

// One PW leaf: the leaf box, its level, its ancestor at the cutoff level,
// its cell (jx, jy) inside that ancestor at the leaf's level, and the scaled
// offset (c_leaf - c_anc)/sqrt(delta0).
typedef short v4s __attribute__((ext_vector_type(4)));  // four bf16 bit patterns (MFMA operand)

struct PwLeaf {
    int box;
    int level;
    int anc;
    int jx;
    int jy;
    int pad;
    double dx, dy;
};

// A group of PW leaves sharing one ancestor; its sum goes to slot `slot` of the
// expansion buffer (the ancestor's own mp slot when it is its only chunk).
struct UpChunk {
    int anc;
    int first;  // into the PwLeaf array
    int count;
    int slot;
};

// One near-field interaction: source box and table offsets (k = -6..6).
struct NearPair {
    int src;
    int ix;
    int iy;
    int pad;
};

// The same interaction seen from the source: target box, target level and
// the table offsets (the adjoint of the near field runs over these).
struct NearPairT {
    int tgt;
    int lev;
    int ix;
    int iy;
};

// Tree node for locating targets: centre, 0-based children in the
// reference's order (mm, pm, mp, pp; -1 for a leaf) and level.
struct TreeNode {
    double cx, cy;
    int child[4];
    int level;
    int pad;
};

// leaves per column group of the downward passes (leaf order > 8), and in
// the short groups (fewer than 2048 long ones; the split-FP32 pass)
constexpr int kColLeaves = 8;
constexpr int kColLeavesShort = 4;
// waves (chunks) of one upward workgroup that sum the chunks of a cutoff box
constexpr int kUpReduceWaves = 4;
// below this many (chunk, y tile) waves the upward pass puts each density on
// its own workgroups
constexpr int kUpSplitWaves = 2048;
// quads (2 x 2 leaf blocks) per column-pair group of the downward pass (n <= 8)
constexpr int kQuadsPerGroup = 4;
// ints per translation group in patch form: 16 cells, 4 target masks
constexpr int kPatchInts = 20;

struct Plan {
    // parameters
    int nd = 1, ipoly = 0, iperiod_in = 1, iperiod_eff = 1, n = 8, npbox = 64, ifpgh = 1;
    int precision = BFGT_PREC_FP64;
    double delta = 0, eps = 0, delta0 = 0, pmax = 0;
    int nlevels = 0, nboxes = 0, nboxes_cap = 0;
    int npwlevel = 0, npw = 0, H = 0, nexp = 0, nlevstart = -1, nlevend = -1;
    bool has_pw = false, has_near = false;
    int nleafbox = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double out_scale = 1.0;  // factor on the outputs of the current execution
    // 1: the evaluation adds into pot/grad/hess; 0: it writes the PW leaves'
    // outputs (an overwriting execution, which first clears the other leaves)
    int accum = 1;
    int nnonpw = 0;            // leaves that neither a PW evaluation nor the near field writes
    int* d_nonpw = nullptr;    // their boxes
    int* d_near_fresh = nullptr;  // per near-field target: 1 if no PW evaluation writes it
    // three-slice bf16 passes: phased, sliced tables per (level, cell)
    int b3nkb = 0;  // 16-long blocks of x frequencies of the downward x contraction
    v4s *d_b3dx = nullptr, *d_b3dy = nullptr, *d_b3uy = nullptr, *d_b3ux = nullptr;
    int device = 0;

    // host copies used for scheduling
    std::vector<double> boxsize, ts, ws;

    // device tables
    double* d_ts = nullptr;       // npw
    double* d_ws = nullptr;       // npw
    double* d_boxsize = nullptr;  // nlevels+1
    double* d_nodes = nullptr;    // node rules and U matrices (see tables)
    cplx* d_tpoly = nullptr;      // [level][k][m]
    cplx* d_tpw = nullptr;        // [level][type][k][i]
    double* d_tloc = nullptr;     // [level][13][type][j][m], masked
    int* d_indloc = nullptr;      // [level][13][n+1][2]
    cplx* d_wshift = nullptr;     // [9][nexp]
    cplx* d_phase = nullptr;      // leaf->ancestor phases, [levbase[l] + j*pstride + k], zero for k >= npw
    int pstride = 0;              // frequency stride of the phase and padded tables (>= npw)
    int* d_levbase = nullptr;     // [nlevels+1]
    std::vector<int> levbase;     // host copy (kept alive for the asynchronous upload)
    size_t nphase = 0;
    std::vector<int> h_indloc;    // copy of d_indloc

    // schedules
    int npwleaf = 0, nchunks = 0, nanc = 0, nm2l = 0;
    PwLeaf* d_pwleaf = nullptr;
    int npair = 0;                 // PW leaves paired by (cutoff box, level) for n <= 8
    int2* d_pwpair = nullptr;      // (leaf, partner leaf or -1)
    int ncolgrp = 0;               // runs of PW leaves with one cutoff box, level and x cell
    long long nxcol_up = 0;        // column pieces of the upward chunks (x contractions)
    int2* d_upcols = nullptr;      // column pieces (first leaf, count), chunk by chunk
    int4* d_upcc = nullptr;        // per chunk: (first column piece, count, slot, 1)
    // the same with the chunks of one cutoff box summed within each upward
    // workgroup (kUpReduceWaves consecutive chunks): (first column piece,
    // count, destination slot, chunks in the piece at its first chunk else 0)
    int4* d_upccb = nullptr;
    int nmergeb = 0;               // the merge of those pieces (as d_merge)
    int4* d_mergeb = nullptr;
    long long merge_reads_b = 0;
    // n <= 8 upward pass: per chunk (first segment, count, destination,
    // run) for either merge; segments (first quad, count, first leaf of
    // column piece A, of B or -1) pair a chunk's column pieces of one level,
    // quads as for the downward pass
    int4* d_upqcc = nullptr;
    int4* d_upqccb = nullptr;
    int4* d_upsegs = nullptr;
    int4* d_upquads = nullptr;
    long long nxcol_upq = 0;       // column pieces in those segments (x contractions of n points)
    bool merge_pieces = false;     // the last upward pass summed chunks per workgroup
    long long merge_reads = 0;     // expansions the merge sums
    int2* d_colgrp = nullptr;      // (first leaf, count)
    int ncolgrp_short = 0;         // the same runs in groups of at most kColLeavesShort
    int2* d_colgrp_short = nullptr;
    int nqgroup = 0;               // n <= 8: column pairs evaluated as 2 x 2 leaf blocks
    int4* d_qgroups = nullptr;     // (first quad, count, a leaf of column A, a leaf of column B or -1)
    int4* d_quads = nullptr;       // leaves (A, jyA), (B, jyA), (A, jyB), (B, jyB) or -1
    long long nxcol_quad = 0;      // columns of the groups (x contractions of n points)
    UpChunk* d_chunks = nullptr;
    int* d_anc_box = nullptr;      // ancestor index -> box
    int* d_anc_chunk = nullptr;    // CSR over chunks [nanc+1]
    // merge of chunk partial sums: one entry per cutoff box with several
    // (anc, first slot, count, 0), the slots consecutive past the mp slots
    int nmerge = 0;
    int4* d_merge = nullptr;
    int* d_m2l_ptr = nullptr;      // CSR [nanc+1]
    int2* d_m2l = nullptr;         // (source ancestor, shift index)
    int nnear_t = 0, nnear_p = 0;
    int* d_near_tgt = nullptr;     // target boxes
    int* d_near_lev = nullptr;     // target level
    int* d_near_ptr = nullptr;     // CSR [nnear_t+1]
    NearPair* d_near = nullptr;
    TreeNode* d_tnodes = nullptr;  // [nboxes], for extra targets
    double* d_tnear = nullptr;     // [level][13][3][16][16] zero-padded near tables (n <= 16), [target][source]
    double* d_tnearT = nullptr;    // the same, [source][target]

    // zero-padded PW tables of the matrix-core kernels (16 point rows or
    // columns, pstride frequencies), so that every load in their inner loops
    // is unconditional and in bounds:
    //   tup_f[l][m][k] = T(m, k)                 values -> PW (upward, forward)
    //   tup_a[l][t][m][k] = conj E_t(k, m)       adjoint of the evaluation (upward)
    //   tdn_f[l][t][k][i] = E_t(k, i)            PW -> values (downward, forward)
    //   tdn_a[l][k][i] = conj T(i, k)            adjoint of the transform (downward)
    cplx* d_tup_f = nullptr;
    cplx* d_tup_a = nullptr;
    cplx* d_tdn_f = nullptr;
    cplx* d_tdn_a = nullptr;

    // adjoint (transposed operator): the transposed PW translation lists and
    // the near pairs grouped by source
    bool has_adjoint = false;
    int* d_m2lT_ptr = nullptr;     // CSR [nanc+1]
    int2* d_m2lT = nullptr;        // (target ancestor, shift index)
    // both lists regrouped by m2l_group (1 or 4) consecutive cutoff boxes: CSR
    // [ngroups+1] of (source, one 4-bit shift index per target, 15 = none)
    int m2l_group = 1;
    int* d_m2lg_ptr = nullptr;
    int2* d_m2lg = nullptr;
    int* d_m2lgT_ptr = nullptr;
    int2* d_m2lgT = nullptr;
    int* d_m2lpatch = nullptr;      // groups of four: patch form of the lists (or -2: use the entries)
    int* d_m2lpatchT = nullptr;
    int nnearT_s = 0;
    int* d_nearT_src = nullptr;    // source boxes
    int* d_nearT_ptr = nullptr;    // CSR [nnearT_s+1]
    NearPairT* d_nearT = nullptr;

    // work buffers
    // [nmpslot][nd][nexp]: the cutoff boxes' mp in slots [0, nanc), then the
    // chunk partial sums of boxes with more than one chunk
    cplx* d_mp = nullptr;
    int nmpslot = 0;
    cplx* d_loc = nullptr;      // [nanc][nd][nexp]

    // timing
    hipEvent_t ev[10] = {};
    std::vector<hipEvent_t> prof_ev;  // 6 events per profiled execution
    int prof_max = 0, prof_count = 0;
    int prof_stage = -1;  // -1: every stage, else the one stage whose two events are recorded
    bool profiling = false;
    double t_plan_host = 0;
    double t_tables = 0;

    std::vector<void*> allocations;
};

// The dispatcher deals workgroups round-robin to the 8 XCDs of the chip,
// each with its own L2.  xcd_block maps physical workgroup b of nb to a
// logical index such that every XCD runs one contiguous range of logical
// workgroups (a bijection on [0, nb)), so neighbouring work items -- leaves
// of one cutoff box, neighbouring cutoff boxes -- share an L2.
constexpr int kXcds = 8;
__device__ inline int xcd_block(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// Split-FP32 products (BFGT_PREC_SPLIT32): an FP64 operand x is split into
// hi = fl32(x) and lo = fl32(x - hi), and a*b is formed as hi*hi + hi*lo +
// lo*hi on the FP32 matrix cores, accumulated in FP32.
// Lane layout of v_mfma_f32_16x16x4f32 on gfx950 (measured): A holds A[c][q],
// B holds B[q][c], element r of the accumulator holds D[4q + r][c] -- the
// transpose of the FP64 instruction's row order.  Register r of lane (q, c)
// therefore serves as the B operand of a K-step whose K index is 4q + r, and
// the A operand of that step is loaded at K index 4q + r.
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline v4f mfma32(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

struct Split {
    float h, l;
};
__device__ inline Split split(double x) {
    float h = (float)x;
    return Split{h, (float)(x - (double)h)};
}
__device__ inline Split neg(Split s) { return Split{-s.h, -s.l}; }
// c += a*b, both factors split
__device__ inline v4f mfma3(Split a, Split b, v4f c) {
    c = mfma32(a.l, b.h, c);
    c = mfma32(a.h, b.l, c);
    return mfma32(a.h, b.h, c);
}
// c += a*b, a split, b an FP32 accumulator value
__device__ inline v4f mfma2(Split a, float b, v4f c) {
    c = mfma32(a.l, b, c);
    return mfma32(a.h, b, c);
}
__device__ inline int drow32(int q, int r) { return 4 * q + r; }

// Three-slice bf16 products (BFGT_PREC_BF16X3): a value x is cut into three
// bf16 slices, x = h + m + l + O(2^-26 |x|) (each slice the bf16 nearest to
// what the previous ones leave), and a*b is formed from the six slice
// products of weight above 2^-24 -- hl, lh, mm, mh, hm, hh, smallest first --
// each exact in FP32, on v_mfma_f32_16x16x16_bf16 with FP32 accumulation.
// Lane layout on gfx950 (measured): A holds A[c][4q..4q+3], B holds
// B[4q..4q+3][c], element r of the accumulator holds D[4q + r][c]: the four
// elements of an FP32 accumulator are exactly the B operand of a following
// product whose 16-long K dimension is that accumulator's rows.
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
struct Slices3 {
    v4s h, m, l;
};
// slices of four FP32 values: each slice rounds (to nearest even, the
// packed conversion v_cvt_pk_bf16_f32) what the previous ones leave, and
// those residuals are exact in FP32
__device__ inline Slices3 slice3(float x0, float x1, float x2, float x3) {
    const v4bf h = {(__bf16)x0, (__bf16)x1, (__bf16)x2, (__bf16)x3};
    const float r0 = x0 - (float)h[0], r1 = x1 - (float)h[1], r2 = x2 - (float)h[2], r3 = x3 - (float)h[3];
    const v4bf m = {(__bf16)r0, (__bf16)r1, (__bf16)r2, (__bf16)r3};
    const v4bf l = {(__bf16)(r0 - (float)m[0]), (__bf16)(r1 - (float)m[1]), (__bf16)(r2 - (float)m[2]),
                    (__bf16)(r3 - (float)m[3])};
    return Slices3{__builtin_bit_cast(v4s, h), __builtin_bit_cast(v4s, m), __builtin_bit_cast(v4s, l)};
}
__device__ inline Slices3 slice3(const float (&x)[4]) { return slice3(x[0], x[1], x[2], x[3]); }
// of four FP64 values: their FP32 roundings carry the 2^-24 the slices keep
__device__ inline Slices3 slice3(const double (&x)[4]) {
    return slice3((float)x[0], (float)x[1], (float)x[2], (float)x[3]);
}
__device__ inline v4f mfmab(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// c += a*b over a 16-long K block, both factors sliced
__device__ inline v4f mfma6(const Slices3& a, const Slices3& b, v4f c) {
    c = mfmab(a.h, b.l, c);
    c = mfmab(a.l, b.h, c);
    c = mfmab(a.m, b.m, c);
    c = mfmab(a.m, b.h, c);
    c = mfmab(a.h, b.m, c);
    return mfmab(a.h, b.h, c);
}

// status helpers
int set_error(int code, const std::string& msg);

// The reference's tree arrays seen through itree and iptr (ids 1-based
// inside the arrays, -1 for none); first()/last() give 0-based box ids.
struct TreeView {
    int nlevels = 0, nboxes = 0;
    int *laddr, *ilevel, *iparent, *nchild, *ichild, *ncoll, *coll;
    const double *centers, *boxsize;
    int first(int l) const { return laddr[2 * l] - 1; }
    int last(int l) const { return laddr[2 * l + 1] - 1; }
};

// 1-D rules of a node family on the host (bfgt_tree.hip): nodes, values ->
// coefficients, child/parent interpolation, tail masks, product weights
struct TreeRules {
    int n = 0, ipoly = 0, iptype = 2;
    std::vector<double> x, w;  // ascending nodes on [-1, 1], quadrature weights
    std::vector<double> U;     // values -> coefficients, [k*n + j]
    std::vector<double> P2C;   // [s][i][k] = P_k(x_i/2 -+ 1/2), s = 0 (minus), 1 (plus)
    std::vector<double> C2P;   // [s][i][k] = P_k(2 x_i + 1) (s = 0, i < n/2), P_k(2 x_{n/2+i} - 1) (s = 1)
    std::vector<double> mask;  // tail mask [kx + n*ky] of the iptype norm
    std::vector<double> mask2; // tail mask of the l2 norm (merging)
    std::vector<double> wq;    // product weights [ix + n*iy]
    double rsum = 1.0, rsum2 = 1.0;
};
TreeRules make_rules(int n, int ipoly, int iptype);

// tree lists (bfgt_plan.cpp)
double per_dist(double d, int iperiod, double bs0);
int check_tree(int nboxes_cap, int nlevels, int ltree, int* itree, const int* iptr, const double* centers,
               const double* boxsize, TreeView& t);
void computecoll(TreeView& t, int iperiod);
std::vector<char> pwexp_flags(const TreeView& t, int npwlevel);
std::vector<std::vector<int>> modified_list1(const TreeView& t, int npwlevel, const std::vector<char>& pwexp,
                                             int iperiod);

// tables (bfgt_tables.hip): launched asynchronously, then waited for
int build_tables(Plan& p);
int finish_tables(Plan& p);

// hot path (bfgt_pw.hip, bfgt_near.hip)
int launch_upward(Plan& p, const double* fvals);
int launch_reduce(Plan& p);
int launch_m2l(Plan& p);
int launch_downward(Plan& p, double* pot, double* grad, double* hess);
int launch_near(Plan& p, const double* fvals, double* pot, double* grad, double* hess);
bool near_use_mfma(const Plan& p);
int build_near_padded(Plan& p);

// tree refinement inside the transform (ifnewtree = 1)
size_t down_smem_bytes(int n, int npw, int ntab);
int unfold_local(Plan& p, cplx* out);                                   // bfgt_pw.hip
int launch_pw_eval_list(Plan& p, const PwLeaf* list, int count, const cplx* loc, double* pot, double* grad,
                        double* hess);                                  // bfgt_pw.hip
int launch_near_list(Plan& p, int ntgt, const int* tgt, const int* tlev, const int* ptr, const NearPair* pairs,
                     const double* tloc, int half, const double* fvals, double* pot, double* grad,
                     double* hess);                                     // bfgt_near.hip
int build_loctab_refined(Plan& p, int mref, double* tloc, int* indloc, int* half);  // bfgt_tables.hip

// The tree side of one host-array call with ifnewtree = 1 (bfgt_newtree.hip):
// the caller's tree arrays (copied in, written back by the caller), the
// plan that has just been executed, and its device value arrays, all of
// capacity nbmax boxes.
struct NewTreeIO {
    int nbmax = 0, ltree = 0;
    int* nlevels = nullptr;   // in/out
    int* itree = nullptr;     // in/out, ltree entries
    const int* iptr = nullptr;
    double* centers = nullptr;  // in/out, 2 * nbmax
    const double* boxsize = nullptr;
    const double* fvals = nullptr;  // device
    double *pot = nullptr, *grad = nullptr, *hess = nullptr;  // device, accumulated in place
    double t_refine = 0, t_coarsen = 0, t_rebalance = 0;
    int nnew = 0, ncapped = 0;
    TreeNode* d_nodes = nullptr;  // device TreeNodes of the final tree (for targets); caller frees
};
int newtree_refine(Plan& p, double eps, NewTreeIO& io);

// leaf-node coordinates (bfgt_targets.hip)
int launch_leaf_points(Plan& p, double* xy);

// extra targets (bfgt_targets.hip)
int launch_targets(Plan& p, int ntarg, const double* targs, int ifpghtarg, const double* pot, const double* grad,
                   const double* hess, double* pote, double* grade, double* hesse);

// adjoint (bfgt_pw.hip, bfgt_near.hip)
bool use_mfma(const Plan& p);
int build_padded_tables(Plan& p);
int build_b3_tables(Plan& p);
size_t fold_len(const Plan& p);
int launch_adj_eval(Plan& p, const double* gpot, const double* ggrad, const double* ghess);
int launch_adj_m2l(Plan& p);
int launch_adj_transform(Plan& p, double* fbar);
int launch_near_adj(Plan& p, const double* gpot, const double* ggrad, const double* ghess, double* fbar);

}  // namespace bfgt

#define BFGT_HIP_CHECK(expr)                                                              \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return ::bfgt::set_error(BFGT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)
