// Host side of the MI355X box FGT: argument checks, the tree lists
// (colleagues, PW flags, modified list 1, PW list), the work schedules that
// the kernels consume, and the C ABI declared in include/boxfgt_hip.h.
//
// The lists restate the reference's tree routines (common/tree_routs_nd.f):
// computecoll :106-242, gnd_find_pwexp_boxes :1266-1368,
// gnd_compute_modified_list1 :864-1055, gnd_compute_listpw :1564-1630; the
// cutoff-level and PW-length choice restates bfgt/boxfgt.f:133-167 and
// common/fgtterms.f:72-141.  They are O(nboxes) integer work done once per
// plan; every floating-point table and every transform runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "bfgt_internal.h"

struct bfgt_plan {
    bfgt::Plan p;
};

namespace bfgt {

thread_local std::string g_err;

int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

size_t node_buffer_doubles(int n);
size_t up_smem_bytes(int n, int npw);
size_t down_smem_bytes(int n, int npw, int ntab);
size_t adj_eval_smem_bytes(int n, int npw, int ntab);
size_t near_smem_bytes(int n, int ntab);
size_t loc_buffer_cplx(const Plan& p);
int build_pw_nodes(Plan& p);

namespace {

constexpr int kChunkLeaves = 8;

}  // namespace

double per_dist(double d, int iperiod, double bs0) {
    d = std::fabs(d);
    if (iperiod == 1) d = std::min(d, bs0 - d);
    return d;
}

int check_tree(int nboxes_cap, int nlevels, int ltree, int* itree, const int* iptr, const double* centers,
               const double* boxsize, TreeView& t) {
    if (nlevels < 0 || ltree < 2 * (nlevels + 1)) return set_error(BFGT_ERR_TREE, "ltree too small for laddr");
    for (int k = 0; k < 7; k++)
        if (iptr[k] < 1 || iptr[k] > ltree) return set_error(BFGT_ERR_TREE, "iptr out of range");
    t.nlevels = nlevels;
    t.laddr = itree + iptr[0] - 1;
    t.ilevel = itree + iptr[1] - 1;
    t.iparent = itree + iptr[2] - 1;
    t.nchild = itree + iptr[3] - 1;
    t.ichild = itree + iptr[4] - 1;
    t.ncoll = itree + iptr[5] - 1;
    t.coll = itree + iptr[6] - 1;
    t.centers = centers;
    t.boxsize = boxsize;
    t.nboxes = itree[2 * nlevels + 1];
    if (t.nboxes < 1 || t.nboxes > nboxes_cap) return set_error(BFGT_ERR_TREE, "box count exceeds nboxes");
    if ((long)iptr[6] - 1 + 9L * t.nboxes > ltree) return set_error(BFGT_ERR_TREE, "itree shorter than its lists");
    int prev = 0;
    for (int l = 0; l <= nlevels; l++) {
        if (t.laddr[2 * l] != prev + 1 || t.laddr[2 * l + 1] < t.laddr[2 * l] - 1)
            return set_error(BFGT_ERR_TREE, "laddr is not a level-sorted partition");
        prev = t.laddr[2 * l + 1];
    }
    for (int l = 0; l <= nlevels; l++)
        for (int b = t.first(l); b <= t.last(l); b++) {
            if (t.ilevel[b] != l) return set_error(BFGT_ERR_TREE, "ilevel disagrees with laddr");
            int nc = t.nchild[b];
            if (nc != 0 && nc != 4) return set_error(BFGT_ERR_TREE, "nchild must be 0 or 4");
            for (int c = 0; c < nc; c++) {
                int k = t.ichild[4 * b + c];
                if (k < 1 || k > t.nboxes || t.ilevel[k - 1] != l + 1)
                    return set_error(BFGT_ERR_TREE, "child id out of range");
            }
            if (l > 0) {
                int dad = t.iparent[b];
                if (dad < 1 || dad > t.nboxes) return set_error(BFGT_ERR_TREE, "parent id out of range");
            }
        }
    for (int l = 1; l <= nlevels; l++)
        if (!(boxsize[l] > 0)) return set_error(BFGT_ERR_TREE, "boxsize must be positive");
    return BFGT_OK;
}

// computecoll (common/tree_routs_nd.f:106-242)
void computecoll(TreeView& t, int iperiod) {
    double bs0 = t.boxsize[0];
    for (int i = 0; i < t.nboxes; i++) {
        t.ncoll[i] = 0;
        for (int j = 0; j < 9; j++) t.coll[9 * i + j] = -1;
    }
    t.ncoll[0] = 1;
    t.coll[0] = 1;
    for (int l = 1; l <= t.nlevels; l++)
        for (int b = t.first(l); b <= t.last(l); b++) {
            int dad = t.iparent[b] - 1;
            for (int i = 0; i < t.ncoll[dad]; i++) {
                int jb = t.coll[9 * dad + i] - 1;
                for (int c = 0; c < 4; c++) {
                    int kb = t.ichild[4 * jb + c];
                    if (kb <= 0) continue;
                    kb -= 1;
                    bool nb = true;
                    for (int d = 0; d < 2 && nb; d++)
                        nb = per_dist(t.centers[2 * kb + d] - t.centers[2 * b + d], iperiod, bs0) <=
                             1.05 * t.boxsize[l];
                    if (nb && t.ncoll[b] < 9) t.coll[9 * b + t.ncoll[b]++] = kb + 1;
                }
            }
        }
}

// gnd_find_pwexp_boxes (common/tree_routs_nd.f:1266-1368)
std::vector<char> pwexp_flags(const TreeView& t, int npwlevel) {
    std::vector<char> f(t.nboxes, 0);
    if (npwlevel <= 0) f[0] = 1;
    if (npwlevel > t.nlevels) return f;
    int nc = std::max(npwlevel, 0);
    for (int b = t.first(nc); b <= t.last(nc); b++) {
        if (t.nchild[b] == 0) {
            for (int j = 0; j < t.ncoll[b]; j++)
                if (t.nchild[t.coll[9 * b + j] - 1] > 0) {
                    f[b] = 1;
                    break;
                }
        } else {
            f[b] = 1;
        }
    }
    for (int l = nc + 1; l <= t.nlevels; l++)
        for (int b = t.first(l); b <= t.last(l); b++) f[b] = 1;
    return f;
}

// gnd_compute_modified_list1 (common/tree_routs_nd.f:864-1055), 0-based ids
std::vector<std::vector<int>> modified_list1(const TreeView& t, int npwlevel, const std::vector<char>& pwexp,
                                             int iperiod) {
    std::vector<std::vector<int>> L(t.nboxes);
    double bs0 = t.boxsize[0];
    if (t.nchild[0] == 0) L[0].push_back(0);
    int nlevend = std::min(t.nlevels, npwlevel);
    auto close = [&](int a, int b, double dt) {
        for (int d = 0; d < 2; d++)
            if (per_dist(t.centers[2 * a + d] - t.centers[2 * b + d], iperiod, bs0) >= dt) return false;
        return true;
    };
    for (int l = 1; l <= nlevend; l++)
        for (int b = t.first(l); b <= t.last(l); b++) {
            if (t.nchild[b] != 0) continue;
            int dad = t.iparent[b] - 1;
            for (int i = 0; i < t.ncoll[b]; i++) {
                int jb = t.coll[9 * b + i] - 1;
                if (t.nchild[jb] == 0) {
                    if (pwexp[b] != 1 || jb != b) L[b].push_back(jb);
                } else if (l < npwlevel) {
                    double dt = 1.05 * (t.boxsize[l] + t.boxsize[l + 1]) / 2.0;
                    for (int c = 0; c < 4; c++) {
                        int kb = t.ichild[4 * jb + c] - 1;
                        if (t.nchild[kb] == 0 && close(kb, b, dt)) L[b].push_back(kb);
                    }
                }
            }
            double dt = 1.05 * (t.boxsize[l] + t.boxsize[l - 1]) / 2.0;
            for (int i = 0; i < t.ncoll[dad]; i++) {
                int jb = t.coll[9 * dad + i] - 1;
                if (t.nchild[jb] == 0 && close(jb, b, dt)) L[b].push_back(jb);
            }
        }
    return L;
}

namespace {

// index of the mp->loc shift (gnd_find_pwshift_ind, bfgt/bfgt_pwrouts.f:1848-1892)
int pwshift_index(int iperiod, const double* tc, const double* sc, double bs0, double xmin) {
    int jx[2];
    for (int i = 0; i < 2; i++) {
        long j = std::lround((tc[i] - sc[i]) / xmin);
        if (iperiod == 1) {
            long jp = std::lround((tc[i] - sc[i] - bs0) / xmin);
            long jm = std::lround((tc[i] - sc[i] + bs0) / xmin);
            if (std::labs(j) > std::labs(jp)) j = jp;
            if (std::labs(j) > std::labs(jm)) j = jm;
        }
        jx[i] = (int)j;
    }
    if (std::abs(jx[0]) > 1 || std::abs(jx[1]) > 1) return -1;
    return (jx[0] + 1) + 3 * (jx[1] + 1);
}

// Translation lists regrouped by G consecutive cutoff boxes (G = 4: the four
// children of one parent, since every refined box has four consecutive
// children): one entry per distinct source expansion, holding the shift index
// of its contribution to each target of the group in a 4-bit field (15 =
// none, 4 = the box itself).  A source that reaches one target under two
// shifts (periodic images on a coarse level) gets a second entry.
void group_lists(int nanc, int G, const std::vector<int>& ptr, const std::vector<int2>& list, std::vector<int>& gptr,
                 std::vector<int2>& glist) {
    gptr.assign(1, 0);
    glist.clear();
    for (int g0 = 0; g0 < nanc; g0 += G) {
        const size_t base = glist.size();
        auto add = [&](int src, int slot, int k) {
            const int sh = 4 * slot;
            for (size_t i = base; i < glist.size(); i++)
                if (glist[i].x == src && ((glist[i].y >> sh) & 15) == 15) {
                    glist[i].y = (glist[i].y & ~(15 << sh)) | (k << sh);
                    return;
                }
            glist.push_back(make_int2(src, (0xFFFF & ~(15 << sh)) | (k << sh)));
        };
        for (int t = 0; t < G && g0 + t < nanc; t++) {
            add(g0 + t, t, 4);
            for (int j = ptr[g0 + t]; j < ptr[g0 + t + 1]; j++) add(list[j].x, t, list[j].y);
        }
        gptr.push_back((int)glist.size());
    }
}

// Patch form of a group of four sibling cutoff boxes (for the translation):
// the group's sources on the 4 x 4 patch of cells around the 2 x 2 targets.
// Record (kPatchInts ints per group): [py*4 + px] the mp slot of the source
// at patch cell (px, py) or -1; [16 + t] for target t its cell (tx | ty << 1)
// << 9 | a 9-bit mask of the source offsets (dx, dy) in its list, bit
// (dx+1) + 3 (dy+1).  A source at offset d from its target is shifted by
// conj(ex)^dx conj(ey)^dy, in the forward (code = target - source) and in
// the transposed lists (code = source - target, conjugated shifts) alike.
// sigma: -1 for forward lists, +1 for transposed.  A group whose entries do
// not map one to one onto the patch (not four siblings, periodic images of a
// coarse level, a partial last group) is marked [0] = -2: the kernel then
// uses the entry list.
void patch_lists(int nanc, const std::vector<int>& gptr, const std::vector<int2>& glist, const TreeView& t,
                 int a0, int sigma, std::vector<int>& patch) {
    const int ng = (nanc + 3) / 4;
    patch.assign((size_t)ng * kPatchInts, -1);
    for (int g = 0; g < ng; g++) {
        int* rec = patch.data() + (size_t)g * kPatchInts;
        bool ok = 4 * g + 3 < nanc;
        int tx[4] = {0, 0, 0, 0}, ty[4] = {0, 0, 0, 0};
        if (ok) {
            const int b0 = a0 + 4 * g;
            double cx = t.centers[2 * b0], cy = t.centers[2 * b0 + 1];
            for (int q = 1; q < 4; q++) {
                cx = std::min(cx, t.centers[2 * (b0 + q)]);
                cy = std::min(cy, t.centers[2 * (b0 + q) + 1]);
            }
            int seen = 0;
            for (int q = 0; q < 4; q++) {
                ok = ok && t.iparent[b0 + q] == t.iparent[b0];
                tx[q] = t.centers[2 * (b0 + q)] > cx ? 1 : 0;
                ty[q] = t.centers[2 * (b0 + q) + 1] > cy ? 1 : 0;
                seen |= 1 << (tx[q] + 2 * ty[q]);
            }
            ok = ok && seen == 15;
        }
        int mask[4] = {0, 0, 0, 0};
        for (int j = gptr[g]; ok && j < gptr[g + 1]; j++) {
            for (int q = 0; q < 4 && ok; q++) {
                const int k = (glist[j].y >> (4 * q)) & 15;
                if (k == 15) continue;
                const int dx = sigma * (k % 3 - 1), dy = sigma * (k / 3 - 1);
                const int px = tx[q] + 1 + dx, py = ty[q] + 1 + dy, bit = (dx + 1) + 3 * (dy + 1);
                int& cell = rec[4 * py + px];
                if ((cell >= 0 && cell != glist[j].x) || (mask[q] >> bit & 1)) ok = false;
                cell = glist[j].x;
                mask[q] |= 1 << bit;
            }
        }
        if (!ok) {
            std::fill(rec, rec + kPatchInts, -1);
            rec[0] = -2;
            continue;
        }
        for (int q = 0; q < 4; q++) rec[16 + q] = ((tx[q] | ty[q] << 1) << 9) | mask[q];
    }
}

// near-field table offsets (gnd_find_loctab_ind, bfgt/bfgt_local.f:1729-1765)
void loctab_index(int iperiod, const double* tc, const double* sc, double sbs, double bs0, int* ixy) {
    double bs = sbs / 4;
    for (int i = 0; i < 2; i++) {
        double dx = (tc[i] - sc[i]) / bs;
        if (iperiod == 1) {
            double dp = dx - bs0 / bs, dm = dx + bs0 / bs;
            if (std::fabs(dx) > std::fabs(dp)) dx = dp;
            if (std::fabs(dx) > std::fabs(dm)) dx = dm;
        }
        ixy[i] = (int)dx;
    }
}

template <class T>
int upload(Plan& p, T*& dst, const std::vector<T>& src) {
    dst = nullptr;
    if (src.empty()) return BFGT_OK;
    BFGT_HIP_CHECK(hipMalloc(&dst, sizeof(T) * src.size()));
    p.allocations.push_back(dst);
    BFGT_HIP_CHECK(hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, p.stream));
    return BFGT_OK;
}

template <class T>
int dalloc(Plan& p, T*& dst, size_t count) {
    dst = nullptr;
    if (count == 0) return BFGT_OK;
    BFGT_HIP_CHECK(hipMalloc(&dst, sizeof(T) * count));
    p.allocations.push_back(dst);
    return BFGT_OK;
}

void release(Plan& p) {
    if (p.stream) hipStreamSynchronize(p.stream);
    for (void* a : p.allocations) hipFree(a);
    p.allocations.clear();
    for (auto& e : p.ev)
        if (e) {
            hipEventDestroy(e);
            e = nullptr;
        }
    for (auto& e : p.prof_ev) hipEventDestroy(e);
    p.prof_ev.clear();
    if (p.own_stream && p.stream) hipStreamDestroy(p.stream);
    p.stream = nullptr;
}

}  // namespace

// fgtpwterms (common/fgtterms.f:72-141)
static void pw_terms(double bsize, double delta, double eps, int iperiod, double* pmax, int* npw) {
    if (iperiod == 0) {
        double rmax = bsize / std::sqrt(delta);
        double d = std::sqrt(std::log(1.0 / eps));
        *pmax = 2.0 * d;
        double h = 2 * M_PI / (rmax + d);
        *npw = 2 * ((int)(*pmax / h) + 1);
    } else {
        *pmax = 0.0;
        *npw = 2 * ((int)(bsize * std::sqrt(std::log(1.0 / eps) / delta) / M_PI) + 1) + 1;
    }
}

int plan_create(Plan& p, int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int nboxes,
                int nlevels, int ltree, int* itree, const int* iptr, const double* centers, const double* boxsize,
                int ifpgh, int precision, void* stream) {
    auto t0 = std::chrono::steady_clock::now();
    if (ndim != 2) return set_error(BFGT_ERR_NDIM, "only ndim=2 is supported");
    if (ipoly != 0 && ipoly != 1) return set_error(BFGT_ERR_IPOLY, "ipoly must be 0 or 1");
    if (ifpgh < 1 || ifpgh > 3) return set_error(BFGT_ERR_IFPGH, "ifpgh must be 1, 2 or 3");
    if (norder < 1 || norder > kMaxOrder) return set_error(BFGT_ERR_NORDER, "norder must be in 1..32");
    if (nd < 1 || !(delta > 0) || !(eps > 0) || eps >= 1 || !itree || !iptr || !centers || !boxsize)
        return set_error(BFGT_ERR_ARG, "invalid scalar or null array argument");
    if (iperiod != 0 && iperiod != 1) return set_error(BFGT_ERR_ARG, "iperiod must be 0 or 1");
    if (precision != BFGT_PREC_FP64 && precision != BFGT_PREC_SPLIT32 && precision != BFGT_PREC_BF16X3)
        return set_error(BFGT_ERR_ARG, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(BFGT_ERR_NODEVICE, "no GPU visible");
    TreeView t;
    int rc = check_tree(nboxes, nlevels, ltree, itree, iptr, centers, boxsize, t);
    if (rc) return rc;

    p.nd = nd;
    p.ipoly = ipoly;
    p.iperiod_in = iperiod;
    p.n = norder;
    p.npbox = norder * norder;
    p.ifpgh = ifpgh;
    p.precision = precision;
    p.delta = delta;
    p.eps = eps;
    p.nlevels = nlevels;
    p.nboxes = t.nboxes;
    p.nboxes_cap = nboxes;
    p.boxsize.assign(boxsize, boxsize + nlevels + 1);
    BFGT_HIP_CHECK(hipGetDevice(&p.device));
    p.stream = (hipStream_t)stream;  // NULL: the device's default stream
    p.own_stream = false;
    // stage-timing events without the system-scope fence at record time (its
    // cache writeback would cost every execution a few microseconds per
    // event); results are handed over through stream order, never via them
    for (auto& e : p.ev) BFGT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));

    // cutoff level and PW length (bfgt/boxfgt.f:133-167)
    double dcut = std::sqrt(delta * std::log(1.0 / eps));
    int npwlevel = nlevels + 1;
    for (int i = nlevels; i >= 0; i--)
        if (boxsize[i] >= dcut) {
            npwlevel = i;
            break;
        }
    if (boxsize[nlevels] > dcut) npwlevel = nlevels + 1;
    if (boxsize[0] <= dcut) npwlevel = 0;
    if (npwlevel <= 1 && iperiod == 1) npwlevel = 0;
    int npw = 0;
    double pmax = 0;
    if (npwlevel <= nlevels) {
        if (npwlevel <= 0) pw_terms(boxsize[0], delta, eps, iperiod, &pmax, &npw);
        else pw_terms(2 * boxsize[npwlevel], delta, eps, 0, &pmax, &npw);
    }
    if (npw > kMaxNpw) return set_error(BFGT_ERR_NPW, "plane-wave length exceeds 200");
    p.pmax = pmax;
    p.npw = npw;
    p.H = (npw + 1) / 2;
    p.nexp = npw * p.H;
    // Fourier-series regime (bfgt/boxfgt.f:492-523)
    bool fourier = (npwlevel <= 1 && iperiod == 1);
    p.iperiod_eff = fourier ? 0 : iperiod;
    if (fourier) npwlevel = 0;
    p.delta0 = fourier ? 1.0 : delta;
    p.npwlevel = npwlevel;
    computecoll(t, p.iperiod_eff);
    p.nlevend = nlevels;
    if (npwlevel < nlevels) p.nlevend = npwlevel;
    if (npwlevel == 0 && iperiod == 1) p.nlevend = -1;
    p.nleafbox = 0;
    for (int l = 1; l <= nlevels; l++)
        for (int b = t.first(l); b <= t.last(l); b++)
            if (t.nchild[b] == 0) p.nleafbox++;
    p.has_pw = !(npwlevel >= nlevels && npwlevel > 0) && npw > 0;
    p.nlevstart = p.has_pw ? std::max(npwlevel, 0) : -1;
    std::vector<char> pwexp = pwexp_flags(t, npwlevel);
    auto list1 = modified_list1(t, npwlevel, pwexp, p.iperiod_eff);

    // resource checks for the kernels
    hipDeviceProp_t prop;
    BFGT_HIP_CHECK(hipGetDeviceProperties(&prop, p.device));
    size_t lds_max = prop.sharedMemPerBlock;
    if (p.has_pw && (up_smem_bytes(p.n, npw) > lds_max || down_smem_bytes(p.n, npw, ifpgh) > lds_max))
        return set_error(BFGT_ERR_RESOURCE, "plane-wave expansion too large for the leaf kernels");
    if (near_smem_bytes(p.n, ifpgh) > lds_max || 9 * p.n * p.n * sizeof(double) > lds_max)
        return set_error(BFGT_ERR_RESOURCE, "leaf order too large for the near-field kernels");

    // device tables
    BFGT_HIP_CHECK(hipSetDevice(p.device));
    rc = upload(p, p.d_boxsize, p.boxsize);
    if (rc) return rc;
    if ((rc = dalloc(p, p.d_nodes, node_buffer_doubles(p.n)))) return rc;
    size_t nn = (size_t)p.n * p.n;
    if ((rc = dalloc(p, p.d_tloc, (size_t)(nlevels + 1) * 13 * 3 * nn))) return rc;
    if ((rc = dalloc(p, p.d_indloc, (size_t)(nlevels + 1) * 13 * 2 * (p.n + 1)))) return rc;
    BFGT_HIP_CHECK(hipMemsetAsync(p.d_tloc, 0, sizeof(double) * (nlevels + 1) * 13 * 3 * nn, p.stream));
    BFGT_HIP_CHECK(hipMemsetAsync(p.d_indloc, 0, sizeof(int) * (nlevels + 1) * 13 * 2 * (p.n + 1), p.stream));
    if (p.has_pw) {
        if ((rc = dalloc(p, p.d_ts, npw))) return rc;
        if ((rc = dalloc(p, p.d_ws, npw))) return rc;
        if ((rc = dalloc(p, p.d_tpoly, (size_t)(nlevels + 1) * npw * p.n))) return rc;
        if ((rc = dalloc(p, p.d_tpw, (size_t)(nlevels + 1) * 3 * npw * p.n))) return rc;
        if ((rc = dalloc(p, p.d_wshift, (size_t)9 * p.nexp))) return rc;
        // leaf->ancestor phase tables: 2^(l-L0) cells per axis at level l
        const int H = (npw + 1) / 2;
        p.pstride = use_mfma(p) ? std::max(npw, 16 * ((H + 15) / 16)) : npw;
        std::vector<int>& levbase = p.levbase;
        levbase.assign(nlevels + 1, 0);
        size_t tot = 0;
        for (int l = p.nlevstart; l <= nlevels; l++) {
            levbase[l] = (int)tot;
            tot += ((size_t)1 << (l - p.nlevstart)) * p.pstride;
            if (tot > ((size_t)1 << 24))
                return set_error(BFGT_ERR_RESOURCE, "tree too deep below the cutoff level for the phase tables");
        }
        p.nphase = tot;
        if ((rc = dalloc(p, p.d_phase, tot))) return rc;
        if ((rc = upload(p, p.d_levbase, levbase))) return rc;
        if ((rc = build_pw_nodes(p))) return rc;
    }
    if ((rc = build_tables(p))) return rc;
    if (near_use_mfma(p)) {
        if ((rc = dalloc(p, p.d_tnear, (size_t)(nlevels + 1) * 13 * 3 * 256))) return rc;
        if ((rc = dalloc(p, p.d_tnearT, (size_t)(nlevels + 1) * 13 * 3 * 256))) return rc;
        if ((rc = build_near_padded(p))) return rc;
    }
    if (p.has_pw && use_mfma(p)) {
        size_t per = (size_t)(nlevels + 1) * 16 * p.pstride;
        if ((rc = dalloc(p, p.d_tup_f, per))) return rc;
        if ((rc = dalloc(p, p.d_tup_a, 3 * per))) return rc;
        if ((rc = dalloc(p, p.d_tdn_f, 3 * per))) return rc;
        if ((rc = dalloc(p, p.d_tdn_a, per))) return rc;
        if ((rc = build_padded_tables(p))) return rc;
        if (p.precision == BFGT_PREC_BF16X3) {
            // the phased, sliced tables per (level, cell); a tree too deep
            // below the cutoff level for them runs split-FP32 instead
            const long ncell = (1L << (nlevels - p.nlevstart + 1)) - 1;
            const long bytes = ncell * 6 * 64 * 8 * (long)(2 * ifpgh + 2) * (((npw + 1) / 2 + 15) / 16 + 1);
            if (bytes > (256L << 20)) p.precision = BFGT_PREC_SPLIT32;
            else if ((rc = build_b3_tables(p))) return rc;
        }
    }

    // PW schedule: leaves grouped by their cutoff-level ancestor
    std::vector<PwLeaf> leaves;
    std::vector<UpChunk> chunks;
    std::vector<int> chunk_cols;  // first column piece of each chunk
    std::vector<int2> upcols;
    std::vector<int> chunk_segs;  // n <= 8: first column-pair segment of each chunk
    std::vector<int4> upsegs, upquads;
    std::vector<int> anc_box, anc_chunk;
    std::vector<int> m2l_ptr;
    std::vector<int2> m2l;
    std::vector<int4> merge, mergeb, upccb;
    if (p.has_pw) {
        int L0 = p.nlevstart;
        int a0 = t.first(L0);
        p.nanc = t.last(L0) - a0 + 1;
        std::vector<std::vector<PwLeaf>> by_anc(p.nanc);
        double sq = std::sqrt(p.delta0);
        for (int l = L0; l <= nlevels; l++)
            for (int b = t.first(l); b <= t.last(l); b++) {
                if (!(pwexp[b] == 1 && t.nchild[b] == 0)) continue;
                int a = b;
                while (t.ilevel[a] > L0) a = t.iparent[a] - 1;
                PwLeaf lf;
                lf.box = b;
                lf.level = l;
                lf.anc = a - a0;
                double half = 0.5 * ((1 << (l - L0)) - 1);
                lf.jx = (int)std::lround((centers[2 * b] - centers[2 * a]) / boxsize[l] + half);
                lf.jy = (int)std::lround((centers[2 * b + 1] - centers[2 * a + 1]) / boxsize[l] + half);
                lf.pad = 0;
                if (lf.jx < 0 || lf.jy < 0 || lf.jx >= (1 << (l - L0)) || lf.jy >= (1 << (l - L0)))
                    return set_error(BFGT_ERR_TREE, "leaf centre outside its cutoff-level ancestor");
                lf.dx = (centers[2 * b] - centers[2 * a]) / sq;
                lf.dy = (centers[2 * b + 1] - centers[2 * a + 1]) / sq;
                by_anc[a - a0].push_back(lf);
            }
        // leaves per chunk: 8, or 16 where the upward pass still has at least
        // eight waves per SIMD (fewer chunk sums to write and merge); 16 for
        // n <= 8, where the chunk's columns are evaluated in pairs
        int chunk_leaves = kChunkLeaves;
        if (use_mfma(p) && p.n <= 8) {
            chunk_leaves = 2 * kChunkLeaves;
        } else if (use_mfma(p)) {
            const long nTN = ((p.npw + 1) / 2 + 15) / 16;
            long waves16 = 0;
            for (int a = 0; a < p.nanc; a++) waves16 += nTN * (((long)by_anc[a].size() + 15) / 16);
            if (waves16 >= 8 * 1024) chunk_leaves = 2 * kChunkLeaves;
        }
        anc_chunk.push_back(0);
        for (int a = 0; a < p.nanc; a++) {
            auto& v = by_anc[a];
            // by level, then x cell, then y cell: the leaves of one column
            // (same box, level and jx) are consecutive
            std::stable_sort(v.begin(), v.end(), [](const PwLeaf& x, const PwLeaf& y) {
                if (x.level != y.level) return x.level < y.level;
                if (x.jx != y.jx) return x.jx < y.jx;
                return x.jy < y.jy;
            });
            // chunks of whole columns up to chunk_leaves leaves (a longer
            // column is cut): the upward pass contracts x once per column
            // piece of a chunk
            const size_t cap = (size_t)chunk_leaves;
            for (size_t s = 0; s < v.size();) {
                size_t e = s;
                while (e < v.size()) {
                    size_t ce = e + 1;
                    while (ce < v.size() && v[ce].level == v[e].level && v[ce].jx == v[e].jx) ce++;
                    if (ce - s <= cap) {
                        e = ce;
                        continue;
                    }
                    if (e == s) e = s + cap;
                    break;
                }
                // the chunk's column pieces
                const int cp0 = (int)upcols.size();
                chunk_cols.push_back(cp0);
                for (size_t i = s; i < e;) {
                    size_t j = i + 1;
                    while (j < e && v[j].level == v[i].level && v[j].jx == v[i].jx) j++;
                    upcols.push_back(make_int2((int)(leaves.size() + i), (int)(j - i)));
                    i = j;
                }
                // n <= 8: consecutive column pieces of one level in pairs,
                // as 2 x 2 leaf blocks (quads, as for the downward pass)
                chunk_segs.push_back((int)upsegs.size());
                if (p.n <= 8) {
                    const int base = (int)leaves.size(), cp1 = (int)upcols.size();
                    for (int k = cp0; k < cp1;) {
                        const int2 ca = upcols[k];
                        const bool two = k + 1 < cp1 && v[upcols[k + 1].x - base].level == v[ca.x - base].level;
                        const int2 cb = two ? upcols[k + 1] : make_int2(0, 0);
                        std::vector<int3> rows;  // (jy, leaf in A or -1, leaf in B or -1)
                        int ia = 0, ib = 0;
                        while (ia < ca.y || ib < cb.y) {
                            const int ja = ia < ca.y ? v[ca.x - base + ia].jy : INT32_MAX;
                            const int jb = ib < cb.y ? v[cb.x - base + ib].jy : INT32_MAX;
                            const int jy = std::min(ja, jb);
                            rows.push_back(make_int3(jy, ja == jy ? ca.x + ia++ : -1, jb == jy ? cb.x + ib++ : -1));
                        }
                        const int q0 = (int)upquads.size();
                        for (size_t r = 0; r < rows.size(); r += 2) {
                            const int3 ra = rows[r], rb = r + 1 < rows.size() ? rows[r + 1] : make_int3(0, -1, -1);
                            upquads.push_back(make_int4(ra.y, ra.z, rb.y, rb.z));
                        }
                        upsegs.push_back(make_int4(q0, (int)upquads.size() - q0, ca.x, two ? cb.x : -1));
                        p.nxcol_upq += two ? 2 : 1;
                        k += two ? 2 : 1;
                    }
                }
                UpChunk c;
                c.anc = a;
                c.first = (int)(leaves.size() + s);
                c.count = (int)(e - s);
                c.slot = (int)chunks.size();
                chunks.push_back(c);
                s = e;
            }
            leaves.insert(leaves.end(), v.begin(), v.end());
            anc_chunk.push_back((int)chunks.size());
            anc_box.push_back(a0 + a);
        }
        // chunk output slots of the expansion buffer: the only chunk of a
        // cutoff box is written straight into the box's mp slot (slots
        // [0, nanc)); the chunks of a box with several get slots from nanc on
        // and are merged.  A box with no chunk keeps the zero mp written when
        // the plan is built.
        int nslot = p.nanc;
        for (int a = 0; a < p.nanc; a++) {
            int c0 = anc_chunk[a], c1 = anc_chunk[a + 1];
            if (c1 - c0 == 1) chunks[c0].slot = a;
            else
                for (int c = c0; c < c1; c++) chunks[c].slot = nslot++;
        }
        p.nmpslot = nslot;
        p.nxcol_up = (long long)upcols.size();
        // merge schedule of the chunk partial sums (consecutive slots)
        for (int a = 0; a < p.nanc; a++) {
            int c0 = anc_chunk[a], c1 = anc_chunk[a + 1], nc = c1 - c0;
            if (nc >= 2) merge.push_back(make_int4(a, chunks[c0].slot, nc, 0));
        }
        // The same merge when the upward pass sums, inside each workgroup of
        // four consecutive chunks, the chunks of one cutoff box (a "piece")
        // before writing: a box whose chunks form one piece is written
        // directly; the pieces of the others get consecutive slots.  Per
        // chunk: (first column piece, count, destination slot, chunks in its
        // piece if it is the piece's first chunk, else 0).
        upccb.resize(chunks.size());
        int nslotb = p.nanc;
        for (int a = 0; a < p.nanc; a++) {
            const int c0 = anc_chunk[a], c1 = anc_chunk[a + 1];
            std::vector<int2> pieces;  // (first chunk, count)
            for (int c = c0; c < c1;) {
                int e = c + 1;
                while (e < c1 && e / kUpReduceWaves == c / kUpReduceWaves) e++;
                pieces.push_back(make_int2(c, e - c));
                c = e;
            }
            const int np = (int)pieces.size(), first = np == 1 ? a : nslotb;
            for (int k = 0; k < np; k++)
                for (int c = pieces[k].x; c < pieces[k].x + pieces[k].y; c++) {
                    const int cf = chunk_cols[c];
                    const int cl = c + 1 < (int)chunk_cols.size() ? chunk_cols[c + 1] : (int)upcols.size();
                    upccb[c] = make_int4(cf, cl - cf, np == 1 ? a : first + k, c == pieces[k].x ? pieces[k].y : 0);
                }
            if (np <= 1) continue;  // written directly, or no PW leaves (mp stays zero)
            nslotb += np;
            mergeb.push_back(make_int4(a, first, np, 0));
        }
        p.nmpslot = std::max(p.nmpslot, nslotb);
        // the upward kernels' LDS holds the sums of waves 1.. of a workgroup
        // only: the first chunk of every workgroup must start a piece
        for (size_t c = 0; c < upccb.size(); c += kUpReduceWaves)
            if (upccb[c].w < 1) return set_error(BFGT_ERR_RESOURCE, "upward piece schedule: workgroup starts mid-piece");
        // PW translation lists at the cutoff level (gnd_compute_listpw)
        double xmin = boxsize[L0];
        m2l_ptr.push_back(0);
        for (int a = 0; a < p.nanc; a++) {
            int b = a0 + a;
            for (int i = 0; i < t.ncoll[b]; i++) {
                int jb = t.coll[9 * b + i] - 1;
                if (jb == b || !(t.nchild[jb] > 0 || t.nchild[b] > 0)) continue;
                int k = pwshift_index(p.iperiod_eff, centers + 2 * b, centers + 2 * jb, boxsize[0], xmin);
                if (k < 0) return set_error(BFGT_ERR_TREE, "PW-list partner is not adjacent");
                m2l.push_back(make_int2(jb - a0, k));
            }
            if ((int)m2l.size() - m2l_ptr.back() > 8)
                return set_error(BFGT_ERR_TREE, "more than 8 PW-list partners (colleague list corrupt)");
            m2l_ptr.push_back((int)m2l.size());
        }
    }
    p.npwleaf = (int)leaves.size();
    // pairs of PW leaves sharing a cutoff box and a level (leaves are grouped by
    // box and sorted by level within it): for n <= 8 one 16-wide MFMA point
    // dimension holds both leaves' points
    std::vector<int2> pairs_pw;
    for (int i = 0; i < p.npwleaf;) {
        const bool two = i + 1 < p.npwleaf && leaves[i + 1].anc == leaves[i].anc && leaves[i + 1].level == leaves[i].level;
        pairs_pw.push_back(make_int2(i, two ? i + 1 : -1));
        i += two ? 2 : 1;
    }
    p.npair = (int)pairs_pw.size();
    // column groups for the downward passes: runs of leaves with the same
    // box, level and x cell, cut to at most kColLeaves leaves where that
    // leaves 2048 groups (one wave each), else to kColLeavesShort; the group's
    // x contraction serves all its leaves and the kernels' registers do not
    // grow with the group.  The split-FP32 pass keeps the short groups.
    std::vector<int2> colgrp, colgrp_short;
    {
        std::vector<int2> cols;
        for (int i = 0; i < p.npwleaf;) {
            int j = i + 1;
            while (j < p.npwleaf && leaves[j].anc == leaves[i].anc && leaves[j].level == leaves[i].level &&
                   leaves[j].jx == leaves[i].jx)
                j++;
            cols.push_back(make_int2(i, j - i));
            i = j;
        }
        auto cut = [&](int cap, std::vector<int2>& out) {
            out.clear();
            for (const int2& c : cols)
                for (int k = 0; k < c.y; k += cap) out.push_back(make_int2(c.x + k, std::min(cap, c.y - k)));
        };
        cut(kColLeavesShort, colgrp_short);
        cut(kColLeaves, colgrp);
        if (colgrp.size() < 2048) colgrp = colgrp_short;
    }
    p.ncolgrp = (int)colgrp.size();
    p.ncolgrp_short = (int)colgrp_short.size();
    // for n <= 8: pairs of columns (same box and level) whose leaves are
    // evaluated as 2 x 2 blocks -- quads (A, jyA), (B, jyA), (A, jyB), (B,
    // jyB) of leaf indices (-1 where no leaf) over consecutive y cells of
    // the union of the two columns' rows -- and groups of at most
    // kQuadsPerGroup quads (first quad, count, a leaf of column A, a leaf of
    // column B or -1), fewer where that is needed for 2048 (group, density)
    // waves
    std::vector<int4> quads, qgroups;
    std::vector<int> qcol_first;  // first quad of each column pair
    if (p.n <= 8) {
        std::vector<int2> cols;  // (first leaf, count) of each column
        for (int i = 0; i < p.npwleaf;) {
            int j = i + 1;
            while (j < p.npwleaf && leaves[j].anc == leaves[i].anc && leaves[j].level == leaves[i].level &&
                   leaves[j].jx == leaves[i].jx)
                j++;
            cols.push_back(make_int2(i, j - i));
            i = j;
        }
        for (size_t c = 0; c < cols.size();) {
            const int2 ca = cols[c];
            const PwLeaf& la = leaves[ca.x];
            const bool pair = c + 1 < cols.size() && leaves[cols[c + 1].x].anc == la.anc &&
                              leaves[cols[c + 1].x].level == la.level;
            const int2 cb = pair ? cols[c + 1] : make_int2(0, 0);
            // rows: the union of the y cells, each with the leaf of A and of B
            std::vector<int3> rows;  // (jy, leaf in A or -1, leaf in B or -1)
            int ia = 0, ib = 0;
            while (ia < ca.y || ib < cb.y) {
                const int ja = ia < ca.y ? leaves[ca.x + ia].jy : INT32_MAX;
                const int jb = ib < cb.y ? leaves[cb.x + ib].jy : INT32_MAX;
                const int jy = std::min(ja, jb);
                rows.push_back(make_int3(jy, ja == jy ? ca.x + ia++ : -1, jb == jy ? cb.x + ib++ : -1));
            }
            qcol_first.push_back((int)quads.size());
            qcol_first.push_back(ca.x);
            qcol_first.push_back(pair ? cb.x : -1);
            for (size_t r = 0; r < rows.size(); r += 2) {
                const int3 ra = rows[r], rb = r + 1 < rows.size() ? rows[r + 1] : make_int3(0, -1, -1);
                quads.push_back(make_int4(ra.y, ra.z, rb.y, rb.z));
            }
            c += pair ? 2 : 1;
        }
        auto cut = [&](int cap) {
            qgroups.clear();
            for (size_t i = 0; i < qcol_first.size(); i += 3) {
                const int q0 = qcol_first[i], q1 = i + 3 < qcol_first.size() ? qcol_first[i + 3] : (int)quads.size();
                for (int k = q0; k < q1; k += cap)
                    qgroups.push_back(make_int4(k, std::min(cap, q1 - k), qcol_first[i + 1], qcol_first[i + 2]));
            }
        };
        for (int cap = kQuadsPerGroup; cap >= 1; cap /= 2) {
            cut(cap);
            if (qgroups.size() * nd >= 2048) break;
        }
        for (const int4& gq : qgroups) p.nxcol_quad += gq.w >= 0 ? 2 : 1;
    }
    p.nqgroup = (int)qgroups.size();
    p.nchunks = (int)chunks.size();
    p.nm2l = (int)m2l.size();
    p.nmerge = (int)merge.size();
    p.merge_reads = 0;
    for (const int4& r : merge) p.merge_reads += r.z;
    p.nmergeb = (int)mergeb.size();
    p.merge_reads_b = 0;
    for (const int4& r : mergeb) p.merge_reads_b += r.z;

    // near-field schedule (bfgt/boxfgt.f:923-1071), pairs with empty tables
    // dropped: needs the sparsity bounds of the tables, which were generated
    // on the device while the PW schedule was built
    if ((rc = finish_tables(p))) return rc;
    std::vector<int> ntgt, nlev, nptr;
    std::vector<NearPair> pairs;
    const int ni = 2 * (p.n + 1);
    auto rows_empty = [&](int lev, int off) {
        const int* ind = p.h_indloc.data() + ((size_t)lev * 13 + (off + 6)) * ni;
        return ind[2 * p.n + 1] - ind[2 * p.n] + 1 == 0;
    };
    bool tree_ok = true;
    auto add_pair = [&](int lev, int src, const int* ixy) {
        if (std::abs(ixy[0]) > 6 || std::abs(ixy[1]) > 6) {
            tree_ok = false;
            return;
        }
        if (rows_empty(lev, ixy[0]) || rows_empty(lev, ixy[1])) return;
        pairs.push_back(NearPair{src, ixy[0], ixy[1], 0});
    };
    if (!(npwlevel == 0 && p.iperiod_eff == 0)) {
        nptr.push_back(0);
        double bs0 = boxsize[0];
        for (int l = 0; l <= p.nlevend; l++)
            for (int b = t.first(l); b <= t.last(l); b++) {
                size_t before = pairs.size();
                for (int jb : list1[b]) {
                    int jl = t.ilevel[jb];
                    double bs = boxsize[jl];
                    int ixy[2];
                    if (p.iperiod_eff == 1 && jl == 1) {
                        loctab_index(0, centers + 2 * b, centers + 2 * jb, bs, bs0, ixy);
                        add_pair(l, jb, ixy);
                        // periodic images of a level-1 source (bfgt/boxfgt.f:951-1053)
                        int im[2], nim = 0;
                        double dxy[2], ct[3][2], bsh = bs * 0.25;
                        for (int i = 0; i < 2; i++) {
                            dxy[i] = centers[2 * jb + i] - centers[2 * b + i];
                            im[i] = std::fabs(dxy[i]) > bsh;
                            nim += im[i];
                        }
                        if (nim > 0) {
                            for (int i = 0; i < 2; i++) {
                                if (dxy[i] < -bsh) ct[0][i] = centers[2 * jb + i] + bs0;
                                else if (dxy[i] > bsh) ct[0][i] = centers[2 * jb + i] - bs0;
                                else ct[0][i] = centers[2 * jb + i];
                            }
                            if (nim == 2) {
                                ct[1][0] = centers[2 * jb];
                                ct[1][1] = ct[0][1];
                                ct[2][0] = ct[0][0];
                                ct[2][1] = centers[2 * jb + 1];
                            }
                            for (int k = 0; k < (1 << nim) - 1; k++) {
                                loctab_index(0, centers + 2 * b, ct[k], bs, bs0, ixy);
                                add_pair(l, jb, ixy);
                            }
                        }
                    } else {
                        loctab_index(p.iperiod_eff, centers + 2 * b, centers + 2 * jb, bs, bs0, ixy);
                        add_pair(l, jb, ixy);
                    }
                }
                if (pairs.size() > before) {
                    ntgt.push_back(b);
                    nlev.push_back(l);
                    nptr.push_back((int)pairs.size());
                }
            }
    }
    if (!tree_ok) return set_error(BFGT_ERR_TREE, "near-field offset outside the 2:1 table set (tree not level-restricted)");
    p.nnear_t = (int)ntgt.size();
    p.nnear_p = (int)pairs.size();
    p.has_near = p.nnear_t > 0;

    // transposed schedules for the adjoint: PW translation lists of the
    // sources, near pairs grouped by source (ascending box, then target order)
    std::vector<int> m2lT_ptr;
    std::vector<int2> m2lT;
    if (p.has_pw) {
        std::vector<std::vector<int2>> byc(p.nanc);
        for (int a = 0; a < p.nanc; a++)
            for (int j = m2l_ptr[a]; j < m2l_ptr[a + 1]; j++) byc[m2l[j].x].push_back(make_int2(a, m2l[j].y));
        m2lT_ptr.push_back(0);
        for (int c = 0; c < p.nanc; c++) {
            if (byc[c].size() > 8) return set_error(BFGT_ERR_TREE, "more than 8 transposed PW-list partners");
            m2lT.insert(m2lT.end(), byc[c].begin(), byc[c].end());
            m2lT_ptr.push_back((int)m2lT.size());
        }
    }
    std::vector<int> nsrc, nsptr;
    std::vector<NearPairT> pairsT;
    {
        std::vector<std::pair<int, NearPairT>> tmp;
        tmp.reserve(pairs.size());
        for (int it = 0; it < p.nnear_t; it++)
            for (int j = nptr[it]; j < nptr[it + 1]; j++)
                tmp.push_back({pairs[j].src, NearPairT{ntgt[it], nlev[it], pairs[j].ix, pairs[j].iy}});
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const std::pair<int, NearPairT>& x, const std::pair<int, NearPairT>& y) {
                             return x.first < y.first;
                         });
        for (size_t j = 0; j < tmp.size(); j++) {
            if (j == 0 || tmp[j].first != tmp[j - 1].first) {
                nsrc.push_back(tmp[j].first);
                nsptr.push_back((int)j);
            }
            pairsT.push_back(tmp[j].second);
        }
        nsptr.push_back((int)pairsT.size());
    }
    p.nnearT_s = (int)nsrc.size();

    if ((rc = upload(p, p.d_pwleaf, leaves))) return rc;
    {
        // for the overwriting execution: leaves that no PW evaluation writes
        // (above the cutoff level, or every leaf of a plan without PW leaves)
        // are written by the near field where they are its targets and
        // cleared beforehand where they are not
        std::vector<char> pwl(t.nboxes, 0), isnt(t.nboxes, 0);
        if (p.has_pw)
            for (int l = p.nlevstart; l <= nlevels; l++)
                for (int b = t.first(l); b <= t.last(l); b++)
                    if (pwexp[b] == 1 && t.nchild[b] == 0) pwl[b] = 1;
        std::vector<int> fresh(ntgt.size());
        for (size_t i = 0; i < ntgt.size(); i++) {
            fresh[i] = pwl[ntgt[i]] ? 0 : 1;
            isnt[ntgt[i]] = 1;
        }
        std::vector<int> nonpw;
        for (int b = 0; b < t.nboxes; b++)
            if (t.nchild[b] == 0 && !pwl[b] && !isnt[b]) nonpw.push_back(b);
        p.nnonpw = (int)nonpw.size();
        if ((rc = upload(p, p.d_nonpw, nonpw))) return rc;
        if ((rc = upload(p, p.d_near_fresh, fresh))) return rc;
    }
    if ((rc = upload(p, p.d_pwpair, pairs_pw))) return rc;
    if ((rc = upload(p, p.d_colgrp, colgrp))) return rc;
    if (p.precision == BFGT_PREC_SPLIT32 && (rc = upload(p, p.d_colgrp_short, colgrp_short))) return rc;
    if ((rc = upload(p, p.d_quads, quads))) return rc;
    if ((rc = upload(p, p.d_qgroups, qgroups))) return rc;
    {
        std::vector<int4> upcc(chunks.size());
        for (size_t c = 0; c < chunks.size(); c++) {
            const int c0 = chunk_cols[c], c1 = c + 1 < chunks.size() ? chunk_cols[c + 1] : (int)upcols.size();
            upcc[c] = make_int4(c0, c1 - c0, chunks[c].slot, 1);
        }
        if ((rc = upload(p, p.d_upcc, upcc))) return rc;
        if ((rc = upload(p, p.d_upccb, upccb))) return rc;
        if (p.n <= 8) {
            // per chunk (first segment, count, destination, run), both merges
            std::vector<int4> sa(chunks.size()), sb(chunks.size());
            for (size_t c = 0; c < chunks.size(); c++) {
                const int s0 = chunk_segs[c], s1 = c + 1 < chunks.size() ? chunk_segs[c + 1] : (int)upsegs.size();
                sa[c] = make_int4(s0, s1 - s0, chunks[c].slot, 1);
                sb[c] = make_int4(s0, s1 - s0, upccb[c].z, upccb[c].w);
            }
            if ((rc = upload(p, p.d_upqcc, sa))) return rc;
            if ((rc = upload(p, p.d_upqccb, sb))) return rc;
            if ((rc = upload(p, p.d_upsegs, upsegs))) return rc;
            if ((rc = upload(p, p.d_upquads, upquads))) return rc;
        }
        if ((rc = upload(p, p.d_upcols, upcols))) return rc;
    }
    if ((rc = upload(p, p.d_chunks, chunks))) return rc;
    if ((rc = upload(p, p.d_anc_box, anc_box))) return rc;
    if ((rc = upload(p, p.d_anc_chunk, anc_chunk))) return rc;
    if ((rc = upload(p, p.d_merge, merge))) return rc;
    if ((rc = upload(p, p.d_mergeb, mergeb))) return rc;
    if ((rc = upload(p, p.d_m2l_ptr, m2l_ptr))) return rc;
    if ((rc = upload(p, p.d_m2l, m2l))) return rc;
    if ((rc = upload(p, p.d_near_tgt, ntgt))) return rc;
    if ((rc = upload(p, p.d_near_lev, nlev))) return rc;
    if ((rc = upload(p, p.d_near_ptr, nptr))) return rc;
    if ((rc = upload(p, p.d_near, pairs))) return rc;
    {
        std::vector<TreeNode> tn(t.nboxes);
        for (int b = 0; b < t.nboxes; b++) {
            TreeNode& q = tn[b];
            q.cx = centers[2 * b];
            q.cy = centers[2 * b + 1];
            for (int c = 0; c < 4; c++) q.child[c] = t.nchild[b] > 0 ? t.ichild[4 * b + c] - 1 : -1;
            q.level = t.ilevel[b];
            q.pad = 0;
        }
        if ((rc = upload(p, p.d_tnodes, tn))) return rc;
    }
    if ((rc = upload(p, p.d_m2lT_ptr, m2lT_ptr))) return rc;
    if ((rc = upload(p, p.d_m2lT, m2lT))) return rc;
    if (p.has_pw) {
        std::vector<int> gptr;
        std::vector<int2> glist;
        // groups of four read each shared neighbour once, but give a quarter
        // of the workgroups: used only where that still fills the chip
        const int H = (p.npw + 1) / 2;
        p.m2l_group = (long)((p.nanc + 3) / 4) * ((p.nd * H * H + 255) / 256) >= 2048 ? 4 : 1;
        std::vector<int> patch;
        group_lists(p.nanc, p.m2l_group, m2l_ptr, m2l, gptr, glist);
        if ((rc = upload(p, p.d_m2lg_ptr, gptr))) return rc;
        if ((rc = upload(p, p.d_m2lg, glist))) return rc;
        if (p.m2l_group == 4) {
            patch_lists(p.nanc, gptr, glist, t, t.first(p.nlevstart), -1, patch);
            if ((rc = upload(p, p.d_m2lpatch, patch))) return rc;
        }
        group_lists(p.nanc, p.m2l_group, m2lT_ptr, m2lT, gptr, glist);
        if ((rc = upload(p, p.d_m2lgT_ptr, gptr))) return rc;
        if ((rc = upload(p, p.d_m2lgT, glist))) return rc;
        if (p.m2l_group == 4) {
            patch_lists(p.nanc, gptr, glist, t, t.first(p.nlevstart), 1, patch);
            if ((rc = upload(p, p.d_m2lpatchT, patch))) return rc;
        }
    }
    if ((rc = upload(p, p.d_nearT_src, nsrc))) return rc;
    if ((rc = upload(p, p.d_nearT_ptr, nsptr))) return rc;
    if ((rc = upload(p, p.d_nearT, pairsT))) return rc;
    // the vector kernels' adjoint of the evaluation stages the evaluation
    // tables of every channel and one channel's cotangents at once
    p.has_adjoint = !p.has_pw || use_mfma(p) || adj_eval_smem_bytes(p.n, p.npw, p.ifpgh) <= lds_max;
    // split-FP32 exists for the matrix-core leaf transforms only; a plan on the
    // vector kernels runs (and reports) FP64
    if (p.has_pw && !use_mfma(p)) p.precision = BFGT_PREC_FP64;
    if (p.has_pw) {
        size_t len = (size_t)nd * p.nexp;
        if ((rc = dalloc(p, p.d_mp, (size_t)p.nmpslot * len))) return rc;
        BFGT_HIP_CHECK(hipMemsetAsync(p.d_mp, 0, (size_t)p.nmpslot * len * sizeof(cplx), p.stream));
        size_t nloc = (size_t)p.nanc * std::max(loc_buffer_cplx(p), 2 * fold_len(p));
        if ((rc = dalloc(p, p.d_loc, nloc))) return rc;
        // the folded layout's padding is never written by the translation: zero it once
        BFGT_HIP_CHECK(hipMemsetAsync(p.d_loc, 0, nloc * sizeof(cplx), p.stream));
    }
    BFGT_HIP_CHECK(hipStreamSynchronize(p.stream));
    auto t1 = std::chrono::steady_clock::now();
    p.t_plan_host = std::chrono::duration<double>(t1 - t0).count() - p.t_tables;
    return BFGT_OK;
}

// profiled stage s (0 leaf values -> PW, 1 merge, 2 translation, 4 PW ->
// values, 5 near field) lies between stage boundaries *lo and *hi
static void stage_marks(int s, int* lo, int* hi) {
    static const int b[6][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 3}, {3, 4}, {4, 5}};
    *lo = b[s][0];
    *hi = b[s][1];
}

// zero pot (grad, hess) of the leaves no PW evaluation writes
__global__ void k_zero_leaves(int nleaf, const int* __restrict__ boxes, int per, int ifpgh, double* __restrict__ pot,
                              double* __restrict__ grad, double* __restrict__ hess) {
    const int nc = ifpgh == 1 ? 1 : (ifpgh == 2 ? 3 : 6);  // doubles per (point, density) over the three arrays
    const long tot = (long)nleaf * per * nc;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e / ((long)per * nc)), r = (int)(e % ((long)per * nc));
        const size_t b = (size_t)boxes[i];
        if (r < per) pot[b * per + r] = 0.0;
        else if (r < 3 * per) grad[b * 2 * per + (r - per)] = 0.0;
        else hess[b * 3 * per + (r - 3 * per)] = 0.0;
    }
}

int plan_execute(Plan& p, const double* fvals, double* pot, double* grad, double* hess, double* timeinfo,
                 double scale, bool overwrite) {
    if (!fvals || !pot || (p.ifpgh >= 2 && !grad) || (p.ifpgh >= 3 && !hess))
        return set_error(BFGT_ERR_ARG, "null device array");
    if (!std::isfinite(scale)) return set_error(BFGT_ERR_ARG, "scale must be finite");
    BFGT_HIP_CHECK(hipSetDevice(p.device));
    p.out_scale = scale;
    p.accum = overwrite ? 0 : 1;
    if (overwrite && p.nnonpw > 0) {
        const int per = p.npbox * p.nd;
        const long tot = (long)p.nnonpw * per * (p.ifpgh == 1 ? 1 : (p.ifpgh == 2 ? 3 : 6));
        const int blocks = (int)std::min<long>((tot + 255) / 256, 4096);
        k_zero_leaves<<<blocks, 256, 0, p.stream>>>(p.nnonpw, p.d_nonpw, per, p.ifpgh, pot, grad, hess);
        BFGT_HIP_CHECK(hipGetLastError());
    }
    int rc;
    // stage boundaries 0..5 (before the upward pass, after it, after the
    // merge, the translation, the downward pass and the near field); a plain
    // execution records none, a profiled one all six or only the two around
    // the profiled stage
    hipEvent_t* ev = timeinfo ? p.ev : nullptr;
    int mark_lo = 0, mark_hi = 5;
    if (!timeinfo && p.profiling && p.prof_count < p.prof_max) {
        ev = p.prof_ev.data() + 6 * p.prof_count++;
        if (p.prof_stage >= 0) stage_marks(p.prof_stage, &mark_lo, &mark_hi);
    }
    auto mark = [&](int i) -> int {
        if (ev && (i == mark_lo || i == mark_hi || (mark_lo == 0 && mark_hi == 5)))
            BFGT_HIP_CHECK(hipEventRecord(ev[i], p.stream));
        return BFGT_OK;
    };
    if ((rc = mark(0))) return rc;
    if (p.has_pw) {
        if ((rc = launch_upward(p, fvals))) return rc;
        if ((rc = mark(1))) return rc;
        if ((rc = launch_reduce(p))) return rc;
        if ((rc = mark(2))) return rc;
        if ((rc = launch_m2l(p))) return rc;
        if ((rc = mark(3))) return rc;
        if ((rc = launch_downward(p, pot, grad, hess))) return rc;
        if ((rc = mark(4))) return rc;
    } else {
        for (int i = 1; i <= 4; i++)
            if ((rc = mark(i))) return rc;
    }
    if ((rc = launch_near(p, fvals, pot, grad, hess))) return rc;
    if ((rc = mark(5))) return rc;
    if (timeinfo) {
        BFGT_HIP_CHECK(hipEventSynchronize(ev[5]));
        float ms[5];
        for (int i = 0; i < 5; i++) BFGT_HIP_CHECK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        for (int i = 0; i < 20; i++) timeinfo[i] = 0;
        timeinfo[1] = ms[0] * 1e-3;  // leaf values -> PW (with the merge folded in)
        timeinfo[2] = ms[1] * 1e-3;  // merge: sum of partial expansions per cutoff box
        timeinfo[3] = ms[2] * 1e-3;  // mp -> loc at the cutoff level
        timeinfo[4] = 0.0;           // split: folded into the evaluation
        timeinfo[5] = ms[3] * 1e-3;  // loc -> pot/grad/hess
        timeinfo[6] = ms[4] * 1e-3;  // near field
    }
    return BFGT_OK;
}

int plan_targets(Plan& p, int ntarg, const double* targs, int ifpghtarg, const double* pot, const double* grad,
                 const double* hess, double* pote, double* grade, double* hesse) {
    if (ntarg < 0) return set_error(BFGT_ERR_ARG, "ntarg must be >= 0");
    if (ifpghtarg < 1 || ifpghtarg > 3) return set_error(BFGT_ERR_TARGETS, "ifpghtarg must be 1, 2 or 3");
    if (ntarg == 0) return BFGT_OK;
    if (!targs || !pot || !pote || (ifpghtarg >= 2 && (!grad || !grade)) || (ifpghtarg >= 3 && (!hess || !hesse)))
        return set_error(BFGT_ERR_ARG, "null device array");
    BFGT_HIP_CHECK(hipSetDevice(p.device));
    return launch_targets(p, ntarg, targs, ifpghtarg, pot, grad, hess, pote, grade, hesse);
}

int plan_adjoint(Plan& p, const double* gpot, const double* ggrad, const double* ghess, double* fbar,
                 double* timeinfo) {
    if (!gpot || !fbar || (p.ifpgh >= 2 && !ggrad) || (p.ifpgh >= 3 && !ghess))
        return set_error(BFGT_ERR_ARG, "null device array");
    if (!p.has_adjoint)
        return set_error(BFGT_ERR_RESOURCE, "the adjoint of this plan's evaluation does not fit the local data share");
    BFGT_HIP_CHECK(hipSetDevice(p.device));
    p.out_scale = 1.0;
    p.accum = 1;  // fbar accumulates
    int rc;
    hipEvent_t* ev = timeinfo ? p.ev : nullptr;  // stage events only when timings are asked for
    auto mark = [&](int i) -> int {
        if (ev) BFGT_HIP_CHECK(hipEventRecord(ev[i], p.stream));
        return BFGT_OK;
    };
    if ((rc = mark(0))) return rc;
    if (p.has_pw) {
        if ((rc = launch_adj_eval(p, gpot, ggrad, ghess))) return rc;
        if ((rc = mark(1))) return rc;
        if ((rc = launch_reduce(p))) return rc;
        if ((rc = mark(2))) return rc;
        if ((rc = launch_adj_m2l(p))) return rc;
        if ((rc = mark(3))) return rc;
        if ((rc = launch_adj_transform(p, fbar))) return rc;
        if ((rc = mark(4))) return rc;
    } else {
        for (int i = 1; i <= 4; i++)
            if ((rc = mark(i))) return rc;
    }
    if ((rc = launch_near_adj(p, gpot, ggrad, ghess, fbar))) return rc;
    if ((rc = mark(5))) return rc;
    if (timeinfo) {
        BFGT_HIP_CHECK(hipEventSynchronize(ev[5]));
        float ms[5];
        for (int i = 0; i < 5; i++) BFGT_HIP_CHECK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        for (int i = 0; i < 20; i++) timeinfo[i] = 0;
        timeinfo[1] = ms[3] * 1e-3;  // transposed leaf transform (PW -> leaf values)
        timeinfo[2] = ms[1] * 1e-3;  // merge of partial cotangent expansions
        timeinfo[3] = ms[2] * 1e-3;  // transposed translation
        timeinfo[5] = ms[0] * 1e-3;  // transposed evaluation (leaf cotangents -> PW)
        timeinfo[6] = ms[4] * 1e-3;  // transposed near field
    }
    return BFGT_OK;
}

// Plans of the host-array entry point.  The drivers call boxfgt on one tree
// several times per time step and over many steps; a plan (schedules, tables,
// device buffers) is kept while the tree arrays it was built from, the
// parameters and the device are the same, and reused.  The key is the part
// of the tree the plan reads (laddr, levels, parents, children of the boxes
// in use, centres, box sizes); the colleague lists are an output of plan
// creation and are written back into the caller's itree on every call.
namespace {
constexpr size_t kHostPlans = 4;  // plans kept (least recently used evicted)

struct HostPlan {
    int dev = -1, nd = 0, ipoly = 0, iperiod = 0, norder = 0, nboxes = 0, nlevels = 0, ifpgh2 = 0, precision = 0;
    double delta = 0, eps = 0;
    std::vector<int> tree;
    std::vector<double> geom;
    std::vector<int> ncoll, coll;  // colleague lists written by plan creation
    bfgt_plan_t plan = nullptr;
    double *dfv = nullptr, *dpot = nullptr, *dgrad = nullptr, *dhess = nullptr;
    size_t cap = 0;  // doubles per density-value array held in dfv / dpot
    unsigned long long used = 0;
};

std::mutex g_host_mu;
std::vector<std::unique_ptr<HostPlan>> g_host_plans;
unsigned long long g_host_clock = 0;

void free_host_plan(HostPlan& h) {
    hipFree(h.dfv);
    hipFree(h.dpot);
    hipFree(h.dgrad);
    hipFree(h.dhess);
    h.dfv = h.dpot = h.dgrad = h.dhess = nullptr;
    h.cap = 0;
    if (h.plan) bfgt_plan_destroy(h.plan);
    h.plan = nullptr;
}

// the key fields of one call (tree parts copied from the caller's arrays,
// whose bounds check_tree has verified)
void host_key(HostPlan& k, const TreeView& t, int nlevels) {
    const int nb = t.nboxes;
    k.tree.clear();
    k.tree.insert(k.tree.end(), t.laddr, t.laddr + 2 * (nlevels + 1));
    k.tree.insert(k.tree.end(), t.ilevel, t.ilevel + nb);
    k.tree.insert(k.tree.end(), t.iparent, t.iparent + nb);
    k.tree.insert(k.tree.end(), t.nchild, t.nchild + nb);
    k.tree.insert(k.tree.end(), t.ichild, t.ichild + 4 * (size_t)nb);
    k.geom.assign(t.centers, t.centers + 2 * (size_t)nb);
    k.geom.insert(k.geom.end(), t.boxsize, t.boxsize + nlevels + 1);
}

bool same_key(const HostPlan& a, const HostPlan& b) {
    return a.dev == b.dev && a.nd == b.nd && a.ipoly == b.ipoly && a.iperiod == b.iperiod && a.norder == b.norder &&
           a.nboxes == b.nboxes && a.nlevels == b.nlevels && a.ifpgh2 == b.ifpgh2 && a.precision == b.precision &&
           std::memcmp(&a.delta, &b.delta, sizeof(double)) == 0 && std::memcmp(&a.eps, &b.eps, sizeof(double)) == 0 &&
           a.tree == b.tree && a.geom.size() == b.geom.size() &&
           std::memcmp(a.geom.data(), b.geom.data(), a.geom.size() * sizeof(double)) == 0;
}

void drop_host_plan(HostPlan* h) {
    for (size_t i = 0; i < g_host_plans.size(); i++)
        if (g_host_plans[i].get() == h) {
            free_host_plan(*h);
            g_host_plans.erase(g_host_plans.begin() + i);
            return;
        }
}

bool host_cache_enabled() {
    const char* v = std::getenv("BFGT_HOST_PLAN_CACHE");
    return !(v && v[0] == '0');
}

// Staging of the host-array entry's copies.  Copies from pageable host
// memory run far below the link rate; the arrays go instead in chunks
// through two pinned buffers, each chunk's host copy split over several host
// threads and overlapping the previous chunk's DMA.  Guarded by g_host_mu.
constexpr size_t kStageBytes = size_t(32) << 20;
constexpr size_t kStageMin = size_t(4) << 20;  // smaller arrays are copied directly
struct Stage {
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t up = nullptr;   // the input copies of a call that builds a plan
    hipEvent_t up_done = nullptr;
    int dev = -1;
};
Stage g_stage;

void stage_free() {
    if (g_stage.up) (void)hipStreamSynchronize(g_stage.up);
    for (int i = 0; i < 2; i++) {
        if (g_stage.buf[i]) (void)hipHostFree(g_stage.buf[i]);
        if (g_stage.ev[i]) (void)hipEventDestroy(g_stage.ev[i]);
        g_stage.buf[i] = nullptr;
        g_stage.ev[i] = nullptr;
    }
    if (g_stage.up_done) (void)hipEventDestroy(g_stage.up_done);
    if (g_stage.up) (void)hipStreamDestroy(g_stage.up);
    g_stage.up_done = nullptr;
    g_stage.up = nullptr;
    g_stage.dev = -1;
}

bool stage_ready(int dev) {
    if (g_stage.dev == dev) return true;
    stage_free();
    for (int i = 0; i < 2; i++)
        if (hipHostMalloc(&g_stage.buf[i], kStageBytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&g_stage.ev[i], hipEventDisableTiming) != hipSuccess) {
            stage_free();
            return false;
        }
    if (hipStreamCreateWithFlags(&g_stage.up, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g_stage.up_done, hipEventDisableTiming) != hipSuccess) {
        stage_free();
        return false;
    }
    g_stage.dev = dev;
    return true;
}

void par_copy(void* dst, const void* src, size_t n) {
    const unsigned nt = std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
    const size_t part = ((n + nt - 1) / nt + 4095) & ~size_t(4095);
    auto copy = [=](size_t off) { std::memcpy((char*)dst + off, (const char*)src + off, std::min(part, n - off)); };
    std::vector<std::thread> th;
    size_t off = part;
    try {
        for (; off < n; off += part) th.emplace_back(copy, off);
    } catch (...) {
        // no thread to be had: this thread copies what is left
        for (; off < n; off += part) copy(off);
    }
    copy(0);
    for (auto& x : th) x.join();
}

// host -> device; k counts the chunks of one call (the buffers alternate)
hipError_t stage_upload(hipStream_t s, void* dev, const void* host, size_t n, int& k) {
    if (n < kStageMin || g_stage.dev < 0) return hipMemcpyAsync(dev, host, n, hipMemcpyHostToDevice, s);
    for (size_t off = 0; off < n; off += kStageBytes, k++) {
        const size_t len = std::min(kStageBytes, n - off);
        const int b = k & 1;
        hipError_t e;
        if (k >= 2 && (e = hipEventSynchronize(g_stage.ev[b])) != hipSuccess) return e;  // its last DMA is done
        par_copy(g_stage.buf[b], (const char*)host + off, len);
        if ((e = hipMemcpyAsync((char*)dev + off, g_stage.buf[b], len, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
        if ((e = hipEventRecord(g_stage.ev[b], s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// device -> host after the stream's work; chunk i + 1 is in flight while
// chunk i is copied out of its buffer
hipError_t stage_download(hipStream_t s, void* host, const void* dev, size_t n) {
    if (n < kStageMin || g_stage.dev < 0) {
        hipError_t e = hipMemcpyAsync(host, dev, n, hipMemcpyDeviceToHost, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    const size_t nc = (n + kStageBytes - 1) / kStageBytes;
    auto issue = [&](size_t i) -> hipError_t {
        const size_t off = i * kStageBytes, len = std::min(kStageBytes, n - off);
        hipError_t e = hipMemcpyAsync(g_stage.buf[i & 1], (const char*)dev + off, len, hipMemcpyDeviceToHost, s);
        return e != hipSuccess ? e : hipEventRecord(g_stage.ev[i & 1], s);
    };
    hipError_t e = issue(0);
    for (size_t i = 0; e == hipSuccess && i < nc; i++) {
        if (i + 1 < nc && (e = issue(i + 1)) != hipSuccess) break;
        if ((e = hipEventSynchronize(g_stage.ev[i & 1])) != hipSuccess) break;
        const size_t off = i * kStageBytes;
        par_copy((char*)host + off, g_stage.buf[i & 1], std::min(kStageBytes, n - off));
    }
    return e;
}
// The inputs of a host-array call on stream s: leaf values and the outputs
// that are accumulated into (channels above ifpgh start from zero).
hipError_t upload_inputs(hipStream_t s, size_t nv, int ifpgh, int ifpgh2, const double* fvals, const double* pot,
                         const double* grad, const double* hess, double* dfv, double* dpot, double* dgrad,
                         double* dhess) {
    int chunk = 0;
    hipError_t e = stage_upload(s, dfv, fvals, nv * sizeof(double), chunk);
    if (e == hipSuccess) e = stage_upload(s, dpot, pot, nv * sizeof(double), chunk);
    if (e == hipSuccess && ifpgh2 >= 2)
        e = ifpgh >= 2 ? stage_upload(s, dgrad, grad, 2 * nv * sizeof(double), chunk)
                       : hipMemsetAsync(dgrad, 0, 2 * nv * sizeof(double), s);
    if (e == hipSuccess && ifpgh2 >= 3)
        e = ifpgh >= 3 ? stage_upload(s, dhess, hess, 3 * nv * sizeof(double), chunk)
                       : hipMemsetAsync(dhess, 0, 3 * nv * sizeof(double), s);
    return e;
}

hipError_t alloc_io(HostPlan& h, size_t nv, int ifpgh2) {
    hipError_t e = hipSuccess;
    if (h.cap >= nv) return e;
    (void)hipFree(h.dfv);
    (void)hipFree(h.dpot);
    (void)hipFree(h.dgrad);
    (void)hipFree(h.dhess);
    h.dfv = h.dpot = h.dgrad = h.dhess = nullptr;
    h.cap = 0;
    if ((e = hipMalloc(&h.dfv, nv * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&h.dpot, nv * sizeof(double))) != hipSuccess) return e;
    if (ifpgh2 >= 2 && (e = hipMalloc(&h.dgrad, 2 * nv * sizeof(double))) != hipSuccess) return e;
    if (ifpgh2 >= 3 && (e = hipMalloc(&h.dhess, 3 * nv * sizeof(double))) != hipSuccess) return e;
    h.cap = nv;
    return e;
}

}  // namespace

// Static workspace of bfgt_plan_execute_noalloc: the cutoff-level mp slots
// and chunk partial sums, then the local expansions, of the plan being
// evaluated, in one file-scope device array instead of the plan's own
// buffers.  Every GPU that runs the library holds its own copy until the
// process ends; its device address is looked up once per device
// (hipGetSymbolAddress) and that address, never the array's host-side name,
// is what the kernels receive.  All calls on one GPU share the array, so
// calls that overlap on one GPU (several host threads, several streams)
// race; the contents left by the previous call are cleared where the
// evaluation needs zeros.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device, null until looked up

int scratch_address(int dev, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(BFGT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return BFGT_OK;
}

// zero fill by a kernel (not a memset node), so that a caller may capture it
__global__ void k_zero_cplx(cplx* __restrict__ a, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = cplx{0.0, 0.0};
}

int zero_cplx(hipStream_t s, cplx* a, size_t n) {
    if (n == 0) return BFGT_OK;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    k_zero_cplx<<<blocks, 256, 0, s>>>(a, n);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

}  // namespace

// plan_execute with the plan's workspace (d_mp, d_loc) replaced, for this
// call only, by the static workspace
int plan_execute_noalloc(Plan& p, const double* fvals, double* pot, double* grad, double* hess, double* timeinfo) {
    if (!fvals || !pot || (p.ifpgh >= 2 && !grad) || (p.ifpgh >= 3 && !hess))
        return set_error(BFGT_ERR_ARG, "null device array");
    if (!p.has_pw) return plan_execute(p, fvals, pot, grad, hess, timeinfo, 1.0, false);  // no workspace used
    const size_t mp_cplx = (size_t)p.nmpslot * p.nd * p.nexp;
    const size_t loc_cplx = (size_t)p.nanc * std::max(loc_buffer_cplx(p), 2 * fold_len(p));
    const size_t mp_bytes = (mp_cplx * sizeof(cplx) + 255) / 256 * 256;
    const size_t need = mp_bytes + loc_cplx * sizeof(cplx);
    if (need > kScratchBytes)
        return set_error(BFGT_ERR_RESOURCE, "the plan's workspace (" + std::to_string(need >> 20) +
                                                " MiB) exceeds the static workspace of " +
                                                std::to_string(kScratchBytes >> 20) + " MiB");
    BFGT_HIP_CHECK(hipSetDevice(p.device));
    void* base = nullptr;
    int rc = scratch_address(p.device, &base);
    if (rc) return rc;
    cplx* mp = static_cast<cplx*>(base);
    cplx* loc = reinterpret_cast<cplx*>(static_cast<char*>(base) + mp_bytes);
    // the mp slots of cutoff boxes without PW leaves are read as zeros, and
    // the padding of the folded local expansions is never written
    if ((rc = zero_cplx(p.stream, mp, (size_t)p.nanc * p.nd * p.nexp))) return rc;
    if ((rc = zero_cplx(p.stream, loc, loc_cplx))) return rc;
    cplx* own_mp = p.d_mp;
    cplx* own_loc = p.d_loc;
    p.d_mp = mp;
    p.d_loc = loc;
    rc = plan_execute(p, fvals, pot, grad, hess, timeinfo, 1.0, false);
    p.d_mp = own_mp;
    p.d_loc = own_loc;
    return rc;
}

}  // namespace bfgt

using bfgt::set_error;

extern "C" {

const char* bfgt_last_error(void) { return bfgt::g_err.c_str(); }

int bfgt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int bfgt_plan_create(bfgt_plan_t* out, int nd, int ndim, double delta, double eps, int ipoly, int iperiod,
                     int norder, int nboxes, int nlevels, int ltree, int* itree, const int* iptr,
                     const double* centers, const double* boxsize, int ifpgh, int precision, void* stream) {
    if (!out) return set_error(BFGT_ERR_ARG, "null plan pointer");
    *out = nullptr;
    std::unique_ptr<bfgt_plan> h;
    int rc;
    try {
        h.reset(new bfgt_plan());
        rc = bfgt::plan_create(h->p, nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels, ltree, itree,
                               iptr, centers, boxsize, ifpgh, precision, stream);
    } catch (const std::exception& e) {  // host allocation: no C++ exception crosses the C ABI
        rc = set_error(BFGT_ERR_RESOURCE, std::string("plan set-up: ") + e.what());
    }
    if (rc) {
        if (h) bfgt::release(h->p);
        return rc;
    }
    *out = h.release();
    return BFGT_OK;
}

int bfgt_plan_execute(bfgt_plan_t plan, const double* d_fvals, double* d_pot, double* d_grad, double* d_hess,
                      double* timeinfo) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    return bfgt::plan_execute(plan->p, d_fvals, d_pot, d_grad, d_hess, timeinfo, 1.0, false);
}

int bfgt_plan_execute_noalloc(bfgt_plan_t plan, const double* d_fvals, double* d_pot, double* d_grad,
                              double* d_hess, double* timeinfo) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    return bfgt::plan_execute_noalloc(plan->p, d_fvals, d_pot, d_grad, d_hess, timeinfo);
}

int bfgt_plan_execute_scaled(bfgt_plan_t plan, double scale, const double* d_fvals, double* d_pot, double* d_grad,
                             double* d_hess, double* timeinfo) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    return bfgt::plan_execute(plan->p, d_fvals, d_pot, d_grad, d_hess, timeinfo, scale, false);
}

int bfgt_plan_execute_overwrite(bfgt_plan_t plan, double scale, const double* d_fvals, double* d_pot,
                                double* d_grad, double* d_hess, double* timeinfo) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    return bfgt::plan_execute(plan->p, d_fvals, d_pot, d_grad, d_hess, timeinfo, scale, true);
}

int bfgt_plan_leaf_points(bfgt_plan_t plan, double* d_xy) {
    if (!plan || !d_xy) return set_error(BFGT_ERR_ARG, "null argument");
    BFGT_HIP_CHECK(hipSetDevice(plan->p.device));
    return bfgt::launch_leaf_points(plan->p, d_xy);
}

int bfgt_plan_eval_targets(bfgt_plan_t plan, int ntarg, const double* d_targs, int ifpghtarg, const double* d_pot,
                           const double* d_grad, const double* d_hess, double* d_pote, double* d_grade,
                           double* d_hesse) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    return bfgt::plan_targets(plan->p, ntarg, d_targs, ifpghtarg, d_pot, d_grad, d_hess, d_pote, d_grade, d_hesse);
}

int bfgt_plan_adjoint(bfgt_plan_t plan, const double* d_gpot, const double* d_ggrad, const double* d_ghess,
                      double* d_fbar, double* timeinfo) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    return bfgt::plan_adjoint(plan->p, d_gpot, d_ggrad, d_ghess, d_fbar, timeinfo);
}

int bfgt_plan_destroy(bfgt_plan_t plan) {
    if (!plan) return BFGT_OK;
    bfgt::release(plan->p);
    delete plan;
    return BFGT_OK;
}

int bfgt_plan_profile_begin(bfgt_plan_t plan, int max_steps) { return bfgt_plan_profile_stage(plan, max_steps, -1); }

int bfgt_plan_profile_stage(bfgt_plan_t plan, int max_steps, int stage) {
    if (!plan || max_steps < 1) return set_error(BFGT_ERR_ARG, "invalid profiling request");
    if (stage != -1 && (stage < 0 || stage > 5 || stage == 3))
        return set_error(BFGT_ERR_ARG, "profiled stage must be -1 (all), 0, 1, 2, 4 or 5");
    bfgt::Plan& p = plan->p;
    BFGT_HIP_CHECK(hipSetDevice(p.device));
    while ((int)p.prof_ev.size() < 6 * max_steps) {
        hipEvent_t e;
        BFGT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        p.prof_ev.push_back(e);
    }
    p.prof_max = max_steps;
    p.prof_count = 0;
    p.prof_stage = stage;
    p.profiling = true;
    return BFGT_OK;
}

int bfgt_plan_profile_end(bfgt_plan_t plan, double* stage_seconds, int* nsteps) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    bfgt::Plan& p = plan->p;
    p.profiling = false;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < p.prof_count; s++) {
        hipEvent_t* ev = p.prof_ev.data() + 6 * s;
        for (int k = 0; k < 6; k++) {
            if (k == 3 || (p.prof_stage >= 0 && k != p.prof_stage)) continue;
            int lo, hi;
            bfgt::stage_marks(k, &lo, &hi);
            BFGT_HIP_CHECK(hipEventSynchronize(ev[hi]));
            float ms;
            BFGT_HIP_CHECK(hipEventElapsedTime(&ms, ev[lo], ev[hi]));
            acc[k] += ms * 1e-3;
        }
    }
    if (stage_seconds) std::memcpy(stage_seconds, acc, sizeof(acc));
    if (nsteps) *nsteps = p.prof_count;
    return BFGT_OK;
}

int bfgt_plan_query(bfgt_plan_t plan, int* iv, double* dv) {
    if (!plan) return set_error(BFGT_ERR_ARG, "null plan");
    const bfgt::Plan& p = plan->p;
    if (iv) {
        int v[16] = {p.npwlevel, p.npw, p.nexp, p.nlevstart, p.nlevend, p.nleafbox, p.npwleaf, p.nchunks,
                     p.nanc,     p.nnear_t, p.nnear_p, p.has_pw, p.has_near, p.iperiod_eff, p.nboxes, p.precision};
        std::memcpy(iv, v, sizeof(v));
    }
    if (dv) {
        double v[8] = {p.delta0, p.pmax, p.t_plan_host, p.t_tables, bfgt::use_mfma(p) ? 1.0 : 0.0, (double)p.nm2l,
                       (double)p.nnearT_s, (double)p.m2l_group};
        std::memcpy(dv, v, sizeof(v));
    }
    return BFGT_OK;
}

int bfgt_plan_work(bfgt_plan_t plan, long long* counts, int ncounts) {
    if (!plan || (!counts && ncounts > 0)) return set_error(BFGT_ERR_ARG, "null argument");
    const bfgt::Plan& p = plan->p;
    const long long L = p.npwleaf;
    const bool mfma = p.has_pw && bfgt::use_mfma(p);
    const bool split = mfma && p.precision == BFGT_PREC_SPLIT32, b3 = mfma && p.precision == BFGT_PREC_BF16X3;
    const bool big = p.n > 8;
    // x contractions of the downward pass: one per column group (n > 8), one
    // per column of a column pair evaluated as 2 x 2 blocks (n <= 8, when
    // there are enough pairs), else one per leaf
    const long long xc_down = big ? p.ncolgrp : ((long long)p.nqgroup * p.nd >= 1024 ? p.nxcol_quad : L);
    // forward passes: split-FP32 shares them as FP64 does, except that its
    // downward pass takes the short column groups (n > 8) and always the
    // column pairs (n <= 8); the three-slice bf16 passes run the column
    // kernels at every order
    const long long fwd_up = !mfma ? L : (b3 || big) ? p.nxcol_up : p.nxcol_upq;
    const long long fwd_down = !mfma ? L : b3 ? p.ncolgrp : split ? (big ? p.ncolgrp_short : p.nxcol_quad) : xc_down;
    // the FP64 and three-slice bf16 upward passes sum chunks per workgroup
    // (at most two x tiles) and then merge workgroup pieces
    const bool pieces = mfma && !split && (p.npw + 1) / 2 <= 32 && p.nmerge > 0;
    long long v[7] = {
        fwd_up,                                                  // forward upward pass
        fwd_down,                                                // forward downward pass
        L,                                                       // PW leaves
        p.has_pw ? (big || p.ifpgh > 1 ? p.nxcol_up : p.nxcol_upq) : L,  // adjoint evaluation
        p.has_pw ? xc_down : L,                                  // adjoint leaf transform
        pieces ? p.merge_reads_b : p.merge_reads,                // expansions read by the merge
        pieces ? (long long)p.nmergeb : (long long)p.nmerge,    // ... and written
    };
    for (int i = 0; i < ncounts && i < 7; i++) counts[i] = v[i];
    for (int i = 7; i < ncounts; i++) counts[i] = 0;
    return BFGT_OK;
}

int bfgt_plan_get_table(bfgt_plan_t plan, int which, int level, void* host_out, size_t nbytes) {
    if (!plan || !host_out) return set_error(BFGT_ERR_ARG, "null argument");
    const bfgt::Plan& p = plan->p;
    if (level < 0 || level > p.nlevels) return set_error(BFGT_ERR_ARG, "level out of range");
    size_t nn = (size_t)p.n * p.n;
    const void* src = nullptr;
    size_t sz = 0;
    switch (which) {
        case 0:
            sz = sizeof(bfgt::cplx) * p.npw * p.n;
            src = p.d_tpoly ? (const void*)(p.d_tpoly + (size_t)level * p.npw * p.n) : nullptr;
            break;
        case 1:
            sz = sizeof(bfgt::cplx) * 3 * p.npw * p.n;
            src = p.d_tpw ? (const void*)(p.d_tpw + (size_t)level * 3 * p.npw * p.n) : nullptr;
            break;
        case 2:
            sz = sizeof(double) * 13 * 3 * nn;
            src = p.d_tloc + (size_t)level * 13 * 3 * nn;
            break;
        case 3:
            sz = sizeof(int) * 13 * 2 * (p.n + 1);
            src = p.d_indloc + (size_t)level * 13 * 2 * (p.n + 1);
            break;
        case 4:
            sz = sizeof(bfgt::cplx) * 9 * p.nexp;
            src = p.d_wshift;
            break;
        case 5:
            if (!p.has_pw || level < p.nlevstart) return set_error(BFGT_ERR_ARG, "no phase table for this level");
            sz = sizeof(bfgt::cplx) * ((size_t)1 << (level - p.nlevstart)) * p.npw;
            if (nbytes < sz) return set_error(BFGT_ERR_ARG, "output buffer too small");
            BFGT_HIP_CHECK(hipMemcpy2D(host_out, sizeof(bfgt::cplx) * p.npw, p.d_phase + p.levbase[level],
                                       sizeof(bfgt::cplx) * p.pstride, sizeof(bfgt::cplx) * p.npw,
                                       (size_t)1 << (level - p.nlevstart), hipMemcpyDeviceToHost));
            return BFGT_OK;
        case 6:
            sz = sizeof(bfgt::cplx) * (size_t)p.nanc * p.nd * p.nexp;
            src = p.d_mp;
            break;
        case 7:
            // MFMA path: mirror-folded (S, D) pairs [anc][d][Mh4][H16] (zero padded); else [anc][d][npw][H]
            sz = sizeof(bfgt::cplx) * (size_t)p.nanc *
                 (bfgt::use_mfma(p) ? 2 * bfgt::fold_len(p) : (size_t)p.nd * p.nexp);
            src = p.d_loc;
            break;
        case 8:
            // near-field schedule: CSR pointer over the target leaves (nnear_t + 1 ints)
            sz = sizeof(int) * (size_t)(p.nnear_t + 1);
            src = p.nnear_t > 0 ? (const void*)p.d_near_ptr : nullptr;
            break;
        default:
            return set_error(BFGT_ERR_ARG, "unknown table");
    }
    if (!src) return set_error(BFGT_ERR_ARG, "table not built for this plan");
    if (nbytes < sz) return set_error(BFGT_ERR_ARG, "output buffer too small");
    BFGT_HIP_CHECK(hipMemcpy(host_out, src, sz, hipMemcpyDeviceToHost));
    return BFGT_OK;
}

static int hip_eval(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int npbox,
                    int nboxes, int nlevels, int ltree, int* itree, const int* iptr, const double* centers,
                    const double* boxsize, const double* fvals, int ifpgh, double* pot, double* grad, double* hess,
                    int ifnewtree, int ntarg, const double* targs, int ifpghtarg, double* pote, double* grade,
                    double* hesse, double* timeinfo, int precision, int* nlevels_io, double* centers_io) {
    auto t0 = std::chrono::steady_clock::now();
    if (ifnewtree != 0 && ifnewtree != 1) return set_error(BFGT_ERR_IFNEWTREE, "ifnewtree must be 0 or 1");
    if (ifnewtree == 1 && (!nlevels_io || !centers_io))
        return set_error(BFGT_ERR_IFNEWTREE, "ifnewtree=1 rewrites the tree: call bfgt_hip_boxfgt");
    if (ifpghtarg < 0 || ifpghtarg > 3) return set_error(BFGT_ERR_TARGETS, "ifpghtarg must be 0..3");
    if (ifpghtarg > 0 && ntarg < 0) return set_error(BFGT_ERR_ARG, "ntarg must be >= 0");
    const bool with_targets = ifpghtarg > 0 && ntarg > 0;
    if (npbox != norder * norder) return set_error(BFGT_ERR_NORDER, "npbox must equal norder**2");
    if (ifpgh < 1 || ifpgh > 3) return set_error(BFGT_ERR_IFPGH, "ifpgh must be 1, 2 or 3");
    if (!fvals || !pot || (ifpgh >= 2 && !grad) || (ifpgh >= 3 && !hess))
        return set_error(BFGT_ERR_ARG, "null output array");
    if (with_targets && (!targs || !pote || (ifpghtarg >= 2 && !grade) || (ifpghtarg >= 3 && !hesse)))
        return set_error(BFGT_ERR_ARG, "null target array");
    if (ndim != 2) return set_error(BFGT_ERR_NDIM, "only ndim=2 is supported");
    if (ipoly != 0 && ipoly != 1) return set_error(BFGT_ERR_IPOLY, "ipoly must be 0 or 1");
    if (norder < 1 || norder > bfgt::kMaxOrder) return set_error(BFGT_ERR_NORDER, "norder must be in 1..32");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(BFGT_ERR_NODEVICE, "no GPU visible");
    // with targets asking for more channels than ifpgh, the transform computes
    // them into zeroed scratch arrays the caller never sees (bfgt/boxfgt.f:207-277)
    const int ifpgh2 = with_targets ? std::max(ifpgh, ifpghtarg) : ifpgh;
    std::lock_guard<std::mutex> lock(bfgt::g_host_mu);
    // the plan for this tree and these parameters: a kept one, or a new one
    bfgt::HostPlan key;
    if (hipGetDevice(&key.dev) != hipSuccess) return set_error(BFGT_ERR_HIP, "hipGetDevice failed");
    key.nd = nd;
    key.ipoly = ipoly;
    key.iperiod = iperiod;
    key.norder = norder;
    key.nboxes = nboxes;
    key.nlevels = nlevels;
    key.ifpgh2 = ifpgh2;
    key.precision = precision;
    key.delta = delta;
    key.eps = eps;
    {
        bfgt::TreeView tv;
        int rc = bfgt::check_tree(nboxes, nlevels, ltree, itree, iptr, centers, boxsize, tv);
        if (rc) return rc;
        bfgt::host_key(key, tv, nlevels);
    }
    const bool keep = bfgt::host_cache_enabled();
    bfgt::HostPlan* hp = nullptr;
    for (auto& h : bfgt::g_host_plans)
        if (bfgt::same_key(*h, key)) hp = h.get();
    const int nused = itree[2 * nlevels + 1];
    const bool reused = hp != nullptr;
    const size_t nv = (size_t)nboxes * npbox * nd;
    double t_plan_host = 0, t_tables = 0;  // of a plan built by this call
    // the pinned staging buffers, set up on the first call on a device
    // (without them the copies go direct)
    const bool staged = bfgt::stage_ready(key.dev);
    if (hp) {
        // the colleague lists plan creation wrote into itree
        std::copy(hp->ncoll.begin(), hp->ncoll.end(), itree + iptr[5] - 1);
        std::copy(hp->coll.begin(), hp->coll.end(), itree + iptr[6] - 1);
    } else {
        // A new plan: the inputs are copied by a second host thread on the
        // staging stream while this one builds the plan (host lists and
        // schedules, device tables); the plan's stream waits for the copies.
        auto io = std::make_unique<bfgt::HostPlan>(std::move(key));
        hipError_t ea = bfgt::alloc_io(*io, nv, ifpgh2);
        if (ea != hipSuccess) {
            bfgt::free_host_plan(*io);
            return set_error(BFGT_ERR_HIP, std::string("device arrays: ") + hipGetErrorString(ea));
        }
        hipError_t eu = hipSuccess;
        std::thread copier;
        auto copy_in = [&]() {
            eu = bfgt::upload_inputs(bfgt::g_stage.up, nv, ifpgh, ifpgh2, fvals, pot, grad, hess, io->dfv, io->dpot,
                                     io->dgrad, io->dhess);
            if (eu == hipSuccess) eu = hipEventRecord(bfgt::g_stage.up_done, bfgt::g_stage.up);
        };
        if (staged) {
            try {
                copier = std::thread(copy_in);
            } catch (...) {
                copy_in();  // no thread to be had: copy first
            }
        }
        hipStream_t stream = nullptr;
        bfgt_plan_t plan = nullptr;
        int rc = BFGT_OK;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
            rc = set_error(BFGT_ERR_HIP, "hipStreamCreateWithFlags failed");
        else
            rc = bfgt_plan_create(&plan, nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels, ltree, itree,
                                  iptr, centers, boxsize, ifpgh2, precision, stream);
        if (copier.joinable()) copier.join();
        if (rc == BFGT_OK && staged && eu != hipSuccess)
            rc = set_error(BFGT_ERR_HIP, std::string("copies in: ") + hipGetErrorString(eu));
        if (rc == BFGT_OK && staged && (eu = hipStreamWaitEvent(stream, bfgt::g_stage.up_done, 0)) != hipSuccess)
            rc = set_error(BFGT_ERR_HIP, std::string("hipStreamWaitEvent: ") + hipGetErrorString(eu));
        if (rc) {
            if (staged) (void)hipStreamSynchronize(bfgt::g_stage.up);
            if (plan) bfgt_plan_destroy(plan);
            else if (stream) (void)hipStreamDestroy(stream);
            bfgt::free_host_plan(*io);
            return rc;
        }
        plan->p.own_stream = true;  // released with the plan
        t_plan_host = plan->p.t_plan_host;
        t_tables = plan->p.t_tables;
        if (bfgt::g_host_plans.size() >= bfgt::kHostPlans) {
            auto lru = std::min_element(bfgt::g_host_plans.begin(), bfgt::g_host_plans.end(),
                                        [](const auto& a, const auto& b) { return a->used < b->used; });
            bfgt::drop_host_plan(lru->get());
        }
        bfgt::g_host_plans.push_back(std::move(io));
        hp = bfgt::g_host_plans.back().get();
        hp->plan = plan;
        hp->ncoll.assign(itree + iptr[5] - 1, itree + iptr[5] - 1 + nused);
        hp->coll.assign(itree + iptr[6] - 1, itree + iptr[6] - 1 + 9 * (size_t)nused);
    }
    hp->used = ++bfgt::g_host_clock;
    bfgt::Plan& p = hp->plan->p;
    const size_t nt = with_targets ? (size_t)ntarg : 0;
    double *dtg = nullptr, *dpe = nullptr, *dge = nullptr, *dhe = nullptr;
    // a failed call frees its target arrays and forgets the plan; a plan is
    // also forgotten after the call when keeping plans is switched off
    auto finish = [&](int code) {
        hipFree(dtg);
        hipFree(dpe);
        hipFree(dge);
        hipFree(dhe);
        if (code != BFGT_OK || !keep) bfgt::drop_host_plan(hp);
        return code;
    };
#define BFGT_EVAL_CHECK(expr)                                                                        \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess) return finish(set_error(BFGT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e))); \
    } while (0)
    BFGT_EVAL_CHECK(bfgt::alloc_io(*hp, nv, ifpgh2));
    double *dfv = hp->dfv, *dpot = hp->dpot, *dgrad = hp->dgrad, *dhess = hp->dhess;
    // the outputs travel both ways: the accumulation (pot += ...) happens on
    // the device.  A new plan's inputs went in while it was built.
    if (reused || !staged)
        BFGT_EVAL_CHECK(bfgt::upload_inputs(p.stream, nv, ifpgh, ifpgh2, fvals, pot, grad, hess, dfv, dpot, dgrad,
                                            dhess));
    if (with_targets) {
        BFGT_EVAL_CHECK(hipMalloc(&dtg, 2 * nt * sizeof(double)));
        BFGT_EVAL_CHECK(hipMalloc(&dpe, nt * nd * sizeof(double)));
        if (ifpghtarg >= 2) BFGT_EVAL_CHECK(hipMalloc(&dge, 2 * nt * nd * sizeof(double)));
        if (ifpghtarg >= 3) BFGT_EVAL_CHECK(hipMalloc(&dhe, 3 * nt * nd * sizeof(double)));
        BFGT_EVAL_CHECK(hipMemcpyAsync(dtg, targs, 2 * nt * sizeof(double), hipMemcpyHostToDevice, p.stream));
    }
    // the copies are timed on their own: wait for them before the kernels
    if (timeinfo) BFGT_EVAL_CHECK(hipStreamSynchronize(p.stream));
    auto t1 = std::chrono::steady_clock::now();
    double ti[20];
    int rc = bfgt::plan_execute(p, dfv, dpot, dgrad, dhess, ti, 1.0, false);
    if (rc) return finish(rc);
    // ifnewtree = 1: the tree is refined, coarsened and rebalanced to
    // resolve pot (bfgt/boxfgt.f:1172-1483); itree, centers and nlevels are
    // rewritten and the values follow the boxes
    bfgt::NewTreeIO ntio;
    if (ifnewtree == 1) {
        ntio.nbmax = nboxes;
        ntio.ltree = ltree;
        ntio.nlevels = nlevels_io;
        ntio.itree = itree;
        ntio.iptr = iptr;
        ntio.centers = centers_io;
        ntio.boxsize = boxsize;
        ntio.fvals = dfv;
        ntio.pot = dpot;
        ntio.grad = dgrad;
        ntio.hess = dhess;
        rc = bfgt::newtree_refine(p, eps, ntio);
        if (rc) {
            hipFree(ntio.d_nodes);
            return finish(rc);
        }
    }
    double t_targets = 0;
    if (with_targets) {
        hipEvent_t e0 = p.ev[6], e1 = p.ev[7];
        BFGT_EVAL_CHECK(hipEventRecord(e0, p.stream));
        bfgt::TreeNode* plan_nodes = p.d_tnodes;
        if (ntio.d_nodes) p.d_tnodes = ntio.d_nodes;  // locate the targets in the new tree
        rc = bfgt::plan_targets(p, ntarg, dtg, ifpghtarg, dpot, dgrad, dhess, dpe, dge, dhe);
        p.d_tnodes = plan_nodes;
        if (rc) {
            hipFree(ntio.d_nodes);
            return finish(rc);
        }
        BFGT_EVAL_CHECK(hipEventRecord(e1, p.stream));
        BFGT_EVAL_CHECK(hipMemcpyAsync(pote, dpe, nt * nd * sizeof(double), hipMemcpyDeviceToHost, p.stream));
        if (ifpghtarg >= 2)
            BFGT_EVAL_CHECK(hipMemcpyAsync(grade, dge, 2 * nt * nd * sizeof(double), hipMemcpyDeviceToHost, p.stream));
        if (ifpghtarg >= 3)
            BFGT_EVAL_CHECK(hipMemcpyAsync(hesse, dhe, 3 * nt * nd * sizeof(double), hipMemcpyDeviceToHost, p.stream));
        BFGT_EVAL_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        BFGT_EVAL_CHECK(hipEventElapsedTime(&ms, e0, e1));
        t_targets = ms * 1e-3;
    }
    hipFree(ntio.d_nodes);
    auto t2 = std::chrono::steady_clock::now();
    BFGT_EVAL_CHECK(bfgt::stage_download(p.stream, pot, dpot, nv * sizeof(double)));
    if (ifpgh >= 2) BFGT_EVAL_CHECK(bfgt::stage_download(p.stream, grad, dgrad, 2 * nv * sizeof(double)));
    if (ifpgh >= 3) BFGT_EVAL_CHECK(bfgt::stage_download(p.stream, hess, dhess, 3 * nv * sizeof(double)));
    BFGT_EVAL_CHECK(hipStreamSynchronize(p.stream));
#undef BFGT_EVAL_CHECK
    if (timeinfo) {
        auto t3 = std::chrono::steady_clock::now();
        for (int i = 0; i < 20; i++) timeinfo[i] = ti[i];
        timeinfo[0] = std::chrono::duration<double>(t1 - t0).count();
        timeinfo[11] = t_targets;  // extra targets (slot 12 of the reference)
        // slots the reference leaves zero: the set-up split into its parts,
        // and the copies back
        timeinfo[13] = t_plan_host;  // host lists and schedules of a new plan
        timeinfo[14] = t_tables;     // its tables (device)
        timeinfo[15] = std::chrono::duration<double>(t1 - t0).count() - t_plan_host - t_tables;  // copies in, checks
        timeinfo[16] = std::chrono::duration<double>(t3 - t2).count();  // copies out
        timeinfo[17] = reused ? 1.0 : 0.0;  // the call ran on a kept plan
        if (ifnewtree == 1) {
            timeinfo[8] = ntio.t_refine;      // slots 9-11 of the reference: refine, coarsen, rebalance
            timeinfo[9] = ntio.t_coarsen;
            timeinfo[10] = ntio.t_rebalance;
            timeinfo[18] = ntio.nnew;         // boxes the refinement created
            timeinfo[19] = ntio.ncapped;      // leaves at the finest level that asked to be split
        }
    }
    return finish(BFGT_OK);
}

int bfgt_hip_eval(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int npbox,
                  int nboxes, int nlevels, int ltree, int* itree, const int* iptr, const double* centers,
                  const double* boxsize, const double* fvals, int ifpgh, double* pot, double* grad, double* hess,
                  int ifnewtree, int ntarg, const double* targs, int ifpghtarg, double* pote, double* grade,
                  double* hesse, double* timeinfo, int precision) {
    try {
        return hip_eval(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree, itree, iptr,
                        centers, boxsize, fvals, ifpgh, pot, grad, hess, ifnewtree, ntarg, targs, ifpghtarg, pote,
                        grade, hesse, timeinfo, precision, nullptr, nullptr);
    } catch (const std::exception& e) {  // host allocation: no C++ exception crosses the C ABI
        return set_error(BFGT_ERR_RESOURCE, std::string("boxfgt: ") + e.what());
    }
}

int bfgt_hip_boxfgt(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int npbox,
                    int nboxes, int* nlevels, int ltree, int* itree, const int* iptr, double* centers,
                    const double* boxsize, const double* fvals, int ifpgh, double* pot, double* grad, double* hess,
                    int ifnewtree, int ntarg, const double* targs, int ifpghtarg, double* pote, double* grade,
                    double* hesse, double* timeinfo, int precision) {
    if (!nlevels || !centers) return set_error(BFGT_ERR_ARG, "null nlevels or centers");
    try {
        return hip_eval(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, *nlevels, ltree, itree, iptr,
                        centers, boxsize, fvals, ifpgh, pot, grad, hess, ifnewtree, ntarg, targs, ifpghtarg, pote,
                        grade, hesse, timeinfo, precision, nlevels, centers);
    } catch (const std::exception& e) {
        return set_error(BFGT_ERR_RESOURCE, std::string("boxfgt: ") + e.what());
    }
}

int bfgt_hip_release_plans(void) {
    std::lock_guard<std::mutex> lock(bfgt::g_host_mu);
    for (auto& h : bfgt::g_host_plans) bfgt::free_host_plan(*h);
    bfgt::g_host_plans.clear();
    bfgt::stage_free();
    return BFGT_OK;
}

}  // extern "C"
