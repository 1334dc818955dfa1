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
// Expansion layout on the device: [slot][density][k1][k2] complex, k1 < npw
// the x frequency, k2 < H = (npw+1)/2 the retained half of the y frequencies
// (the reference's pwexp(npw, H, nd), bfgt/bfgt_pwrouts.f:98-147, with the
// two frequency axes swapped so that tiles of y frequencies are contiguous).
//
// Two implementations: MFMA kernels (v_mfma_f64_16x16x4f64) for leaf orders
// n <= 16 and expansions of at most 128 plane waves, and vector-FMA kernels
// for the remaining sizes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bfgt_internal.h"

namespace bfgt {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline cplx cmul(cplx a, cplx b) { return cplx{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline cplx cfma(cplx a, cplx b, cplx c) {
    return cplx{fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
__device__ inline cplx cexpi(double t) {
    double s, c;
    sincos(t, &s, &c);
    return cplx{c, s};
}
__device__ inline v4d mfma(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

// ---------------------------------------------------------------------------
// MFMA kernels
// ---------------------------------------------------------------------------
//
// Lane layout of v_mfma_f64_16x16x4f64 on gfx950 (lane l, q = l/16,
// c = l%16), measured on the device: A (16x4) holds A[c][q], B (4x16) holds
// B[q][c], and element r of the 16x16 accumulator holds D[4r + q][c].
__device__ inline int drow(int q, int r) { return 4 * r + q; }

constexpr int kUpMaxTM = 4;  // x-frequency tiles of 16 (ceil(npw/2) <= 64)

// A leaf's y-frequency MFMA accumulator holds, in element r of lane (q, c),
// row 4r + q -- exactly the B operand of K-step r of a following MFMA whose
// K index is that row.  Both passes below chain two GEMMs through registers
// this way, with no LDS staging.

// Output channels of the evaluation, as (x table, y table) pairs: pot (0,0),
// d/dx (1,0), d/dy (0,1), d2/dx2 (2,0), d2/dxdy (1,1), d2/dy2 (0,2), where
// table 0, 1, 2 is the value, first and second derivative table.
constexpr int nchan(int ntab) { return ntab == 1 ? 1 : (ntab == 2 ? 3 : 6); }
constexpr int chan_x(int o) { return o == 1 || o == 4 ? 1 : (o == 3 ? 2 : 0); }
constexpr int chan_y(int o) { return o == 2 || o == 4 ? 1 : (o == 5 ? 2 : 0); }

// Leaf arrays read by the upward pass: the leaf values (forward), or the
// cotangents of pot, grad and hess (adjoint of the evaluation).
struct LeafIn {
    const double* pot;   // [box][npbox][nd]
    const double* grad;  // [box][npbox][2][nd]
    const double* hess;  // [box][npbox][3][nd]
};

__device__ inline double leaf_in(const LeafIn& in, int o, size_t pb, int nd, int d) {
    switch (o) {
        case 0: return in.pot[pb * nd + d];
        case 1: return in.grad[(pb * 2 + 0) * nd + d];
        case 2: return in.grad[(pb * 2 + 1) * nd + d];
        case 3: return in.hess[(pb * 3 + 0) * nd + d];
        case 4: return in.hess[(pb * 3 + 1) * nd + d];
        default: return in.hess[(pb * 3 + 2) * nd + d];
    }
}

// The matrix-core kernels read zero-padded tables (16 point rows/columns,
// ps >= npw frequencies, zero beyond npw), zero-padded phases and a
// zero-padded folded expansion, so their inner loops have no bounds checks:
// every load is unconditional and in bounds, and padding rows or columns
// either contribute exact zeros or land in outputs that are not stored.

// Buffer accesses: a wave-uniform base in a buffer resource (scalar
// registers), a 32-bit lane offset in one vector register and a uniform byte
// offset in a scalar register.  The per-step and per-table parts of an
// address then cost no vector registers, and the compiler cannot fold them
// into 64-bit per-lane addresses that stay live across a loop.  Every base
// here is the start of one table, phase level, folded expansion or leaf, so
// the offsets stay far below the 4 GB range of the resource.
// The resource is raw (stride 0), 4 GB - 1 bytes long, with the 32-bit
// data format in its last word (the format untyped loads and stores expect).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ inline rsrc_t rsrc(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, 0x00020000);
}
__device__ inline cplx bload_c(rsrc_t r, unsigned voff, int soff) {
    return __builtin_bit_cast(cplx, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ inline double bload_d(rsrc_t r, unsigned voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ inline void bstore_d(double v, rsrc_t r, unsigned voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
                                          voff, soff, 0);
}

// Upward pass.  One wave per (chunk, y-frequency tile tN): for the leaves of
// the chunk
//   G_v  = sum_o F_o (T_y(o) diag py w)    (rows m1 < 16, columns k2 of the tile)
//   out += sum_v (T_v diag px)^T G_v       with T(m, npw-1-k) = conj T(m, k), so
// that only ceil(npw/2) x frequencies are multiplied; the accumulators AC,
// BD, AD, BC give both halves of the x spectrum at the end.  T_t(m, k) is
// tab[level*lstride + t*16*ps + m*ps + k]; o runs over the channels of NTAB
// tables (chan_x, chan_y) and v over the x tables.  The forward transform
// (bfgt/bfgt_pwrouts.f:98-147) is NTAB = 1 with the values-to-PW table and
// w = 1; the adjoint of the evaluation reads the cotangents with the
// conjugated evaluation tables and w = 1/2 on the middle y frequency of an
// odd expansion.  Leaf values are read at clamped point indices (finite
// values) and meet zero table rows beyond n.  NKS = ceil(n/4),
// NTM = ceil(ceil(npw/2)/16).  Workgroup b serves y tile b % nTN of the
// kUpReduceWaves consecutive chunks b / nTN; with BR the chunks of one
// cutoff box in the workgroup add their sums through LDS and the first of
// them writes (chunk entries (., ., destination slot, run)), else every
// chunk writes its own slot.

// Upward pass for leaf orders n <= 8, one table: the column pieces of a
// chunk in pairs (a segment: column pieces A and B of one level, x cells jxA
// and jxB) and their leaves as 2 x 2 blocks (quads, as in the downward
// pass).  Per quad the y contraction G += F (T diag py) runs with the two
// columns in the two 8-wide halves of G's rows (x points) and the quad's two
// y cells in the two halves of the 16-long K dimension (y points), so one
// product serves four leaves; per segment the x contraction out += (T diag
// px)^T G runs once with the two columns in the two halves of its K
// dimension.  Missing leaves read as zero; ymid_half weights the middle y
// frequency by 1/2 (the adjoint of the evaluation).
template <int NTM, bool BR, bool DS>
__global__ __launch_bounds__(kWave * kUpReduceWaves) void k_up_quad(
    int n, int P, int ps, int nd, int nTN, int nchunks, const double* __restrict__ fvals,
    const PwLeaf* __restrict__ leaves, const int4* __restrict__ chunks, const int4* __restrict__ segs,
    const int4* __restrict__ quads, const cplx* __restrict__ tab, int lstride, int ymid_half,
    const cplx* __restrict__ phase, const int* __restrict__ levbase, cplx* __restrict__ partial) {
    constexpr int NACC = 16 * NTM;  // accumulator doubles per lane
    // the sums of waves 1.. of a workgroup (wave 0 always starts a piece and
    // only reads)
    __shared__ double red[BR ? (kUpReduceWaves - 1) * NACC * kWave : 1];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    // XCD-contiguous: the y tiles of one chunk run on one XCD (one L2), so
    // the chunk's leaf values come from HBM once
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int tN = lb % nTN;
    // DS: one density per workgroup (the y tiles and densities of a chunk on
    // one XCD); else the densities in turn
    const int nds = DS ? nd : 1;
    const int d0 = DS ? (lb / nTN) % nd : 0;
    const int ci = __builtin_amdgcn_readfirstlane((lb / nTN / nds) * kUpReduceWaves + wave);
    const bool valid = ci < nchunks;  // waves past the end still meet the barriers
    const int q = lane >> 4, cc = lane & 15;
    const int half = cc >> 3, xi = min(cc & 7, n - 1);
    const int H = (P + 1) / 2, Mlo = (P + 1) / 2, nexp = P * H, npbox = n * n;
    const int k2 = 16 * tN + cc;
    const double wy = (ymid_half && k2 >= P / 2) ? 0.5 : 1.0;
    const int4 ch = chunks[valid ? ci : 0];
    for (int d = d0; d < (DS ? d0 + 1 : nd); d++) {
        v4d acc[4][NTM];
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int t = 0; t < NTM; t++) acc[p][t] = v4d{0, 0, 0, 0};
        for (int sg = ch.x; valid && sg < ch.x + ch.y; sg++) {
            const int4 seg = segs[sg];
            const PwLeaf la = leaves[seg.z];
            const int jxb = seg.w >= 0 ? leaves[seg.w].jx : la.jx;
            const double wb = seg.w >= 0 ? 1.0 : 0.0;
            const cplx* T = tab + (size_t)la.level * lstride;
            const cplx* ph = phase + levbase[la.level];
            v4d Gr = v4d{0, 0, 0, 0}, Gi = v4d{0, 0, 0, 0};
            for (int k = seg.x; k < seg.x + seg.y; k++) {
                const int4 qd = quads[k];
                // rows of the quad: y cells of (A or B) in row A and in row B
                const int ra = qd.x >= 0 ? qd.x : qd.y, rb = qd.z >= 0 ? qd.z : qd.w;
                const int jya = leaves[ra].jy, jyb = rb >= 0 ? leaves[rb].jy : jya;
                // this lane's leaves (its x half) in rows A and B, zero weight if absent
                const int lfa = half ? qd.y : qd.x, lfb = half ? qd.w : qd.z;
                const size_t pba = (size_t)leaves[lfa >= 0 ? lfa : ra].box * npbox;
                const size_t pbb = (size_t)leaves[lfb >= 0 ? lfb : ra].box * npbox;
                const double wa = lfa >= 0 ? 1.0 : 0.0, wbr = lfb >= 0 ? 1.0 : 0.0;
                cplx pya = ph[jya * ps + k2], pyb = ph[jyb * ps + k2];
                pya.x *= wy;  // upward shift: conjugate of the table
                pya.y *= -wy;
                pyb.x *= wy;
                pyb.y *= -wy;
                if (rb < 0) pyb = cplx{0.0, 0.0};
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int yp = 4 * (s & 1) + q;  // y point of this K index
                    const cplx py = (s >> 1) ? pyb : pya;
                    const cplx ty = cmul(T[min(yp, 15) * ps + k2], py);
                    const double a = ((s >> 1) ? wbr : wa) *
                                     fvals[(((s >> 1) ? pbb : pba) + xi + (size_t)n * min(yp, n - 1)) * nd + d];
                    Gr = mfma(a, ty.x, Gr);
                    Gi = mfma(a, ty.y, Gi);
                }
            }
#pragma unroll
            for (int t = 0; t < NTM; t++) {
                const int k1 = 16 * t + cc;
                cplx pxa = ph[la.jx * ps + k1], pxb = ph[jxb * ps + k1];
                pxa.y = -pxa.y;
                pxb.x *= wb;
                pxb.y *= -wb;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int xp = 4 * (s & 1) + q;  // x point of this K index, column s >> 1
                    const cplx tx = cmul(T[xp * ps + k1], (s >> 1) ? pxb : pxa);
                    acc[0][t] = mfma(tx.x, Gr[s], acc[0][t]);
                    acc[1][t] = mfma(tx.y, Gi[s], acc[1][t]);
                    acc[2][t] = mfma(tx.x, Gi[s], acc[2][t]);
                    acc[3][t] = mfma(tx.y, Gr[s], acc[3][t]);
                }
            }
        }
        if constexpr (BR) {
            // the waves that continue a piece hand their sums to its first wave
            if (valid && ch.w == 0) {
                double* mine = red + (wave - 1) * NACC * kWave;
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int t = 0; t < NTM; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) mine[((p * NTM + t) * 4 + r) * kWave + lane] = acc[p][t][r];
            }
            __syncthreads();
            for (int w = wave + 1; valid && w < wave + ch.w; w++) {
                const double* other = red + (w - 1) * NACC * kWave;
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int t = 0; t < NTM; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) acc[p][t][r] += other[((p * NTM + t) * 4 + r) * kWave + lane];
            }
            if constexpr (!DS) __syncthreads();  // the slots are refilled for the next density
        }
        if (!valid || ch.w == 0) continue;
        cplx* out = partial + ((size_t)ch.z * nd + d) * nexp;
#pragma unroll
        for (int t = 0; t < NTM; t++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int k1 = 16 * t + drow(q, r);
                if (k1 < Mlo && k2 < H) {
                    double AC = acc[0][t][r], BD = acc[1][t][r], AD = acc[2][t][r], BC = acc[3][t][r];
                    out[k1 * H + k2] = cplx{AC - BD, AD + BC};
                    int km = P - 1 - k1;
                    if (km > k1) out[km * H + k2] = cplx{AC + BD, AD - BC};
                }
            }
        }
    }
}

// Upward pass over columns (leaf order > 8, and the adjoint of the
// evaluation): one wave per (chunk, y-frequency tile tN), workgroup b serving
// tile b % nTN of the kUpReduceWaves consecutive chunks b / nTN.  chunks[i] =
// (first column piece, count, destination slot, run) and cols[c] = (first
// leaf, count) of a column piece: PW leaves of one cutoff box, level and x
// cell.  With BR the waves of a workgroup whose chunks belong to one cutoff
// box add their sums through LDS and the first of them (run = their count)
// writes; without BR every chunk writes its own slot (run = 1).  The y
// contractions of a column's leaves are summed in G, LPC leaves per pass (the
// level's table rows are read once for all of them; a pass past the end of
// the column repeats its last leaf with weight zero), and contracted in x
// once:  sum_l (T px)^T F_l (T py_l) = (T px)^T sum_l F_l (T py_l).
template <int NKS, int NTM, int NTAB, int LPC, bool BR, bool DS>
__global__ __launch_bounds__(kWave * kUpReduceWaves) void k_up_col(
    int n, int P, int ps, int nd, int nTN, int nchunks, LeafIn in, const PwLeaf* __restrict__ leaves,
    const int4* __restrict__ chunks, const int2* __restrict__ cols, const cplx* __restrict__ tab, int lstride,
    int ymid_half, const cplx* __restrict__ phase, const int* __restrict__ levbase, cplx* __restrict__ partial) {
    constexpr int NOUT = nchan(NTAB);
    constexpr int NACC = 16 * NTM;  // accumulator doubles per lane
    // the sums of waves 1.. of a workgroup (wave 0 always starts a piece and
    // only reads)
    __shared__ double red[BR ? (kUpReduceWaves - 1) * NACC * kWave : 1];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    // wave-uniform: the schedule is read with scalar loads
    // XCD-contiguous: the y tiles of one chunk run on one XCD (one L2), so
    // the chunk's leaf values come from HBM once
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int tN = lb % nTN;
    // DS: one density per workgroup (the y tiles and densities of a chunk on
    // one XCD); else the densities in turn
    const int nds = DS ? nd : 1;
    const int d0 = DS ? (lb / nTN) % nd : 0;
    const int ci = __builtin_amdgcn_readfirstlane((lb / nTN / nds) * kUpReduceWaves + wave);
    const bool valid = ci < nchunks;  // waves past the end still meet the barriers
    const int q = lane >> 4, cc = lane & 15;
    const int H = (P + 1) / 2, Mlo = (P + 1) / 2, nexp = P * H, npbox = n * n;
    const int tbytes = 16 * ps * (int)sizeof(cplx);  // one table of a level
    const int rowb = ps * (int)sizeof(cplx);         // one point row of a table
    const int k2 = 16 * tN + cc;
    const double wy = (ymid_half && k2 >= P / 2) ? 0.5 : 1.0;
    const int ccl = min(cc, n - 1);
    // lane offsets (bytes): table row q, frequency k2 (y) or cc (x tile);
    // phase frequency k2 or cc; leaf values at the clamped point (cc,
    // m2 = 4s + q) of a density-value array with one component
    const unsigned oTy = (unsigned)(q * ps + k2) * sizeof(cplx);
    const unsigned oTx = (unsigned)(q * ps + cc) * sizeof(cplx);
    const unsigned oPy = (unsigned)k2 * sizeof(cplx);
    const unsigned oPx = (unsigned)cc * sizeof(cplx);
    unsigned oF[NKS];
#pragma unroll
    for (int s = 0; s < NKS; s++) oF[s] = (unsigned)(ccl + n * min(4 * s + q, n - 1)) * (unsigned)(nd * sizeof(double));
    constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
    constexpr int NARR = NTAB;  // leaf arrays read: pot, grad, hess
    const int4 ch = chunks[valid ? ci : 0];
    for (int d = d0; d < (DS ? d0 + 1 : nd); d++) {
        v4d acc[4][NTM];
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int t = 0; t < NTM; t++) acc[p][t] = v4d{0, 0, 0, 0};
        for (int c = ch.x; valid && c < ch.x + ch.y; c++) {
            const int2 col = cols[c];
            const PwLeaf l0 = leaves[col.x];
            const rsrc_t rT = rsrc(tab + (size_t)l0.level * lstride);  // [t][m][k]
            const rsrc_t rP = rsrc(phase + levbase[l0.level]);         // [cell][k]
            v4d Gr[NTAB], Gi[NTAB];
#pragma unroll
            for (int v = 0; v < NTAB; v++) Gr[v] = Gi[v] = v4d{0, 0, 0, 0};
            // y contraction, LPC leaves per pass; the operands of the next
            // K-step (of this pass or of the next pass's first) are loaded
            // before the products of the current one
            struct Pass {
                rsrc_t r[LPC][NARR];
                cplx py[LPC];
            };
            auto pass_at = [&](int li) {
                Pass o;
#pragma unroll
                for (int h = 0; h < LPC; h++) {
                    const PwLeaf lf = leaves[col.x + min(li + h, col.y - 1)];
                    const double wl = li + h < col.y ? 1.0 : 0.0;
#pragma unroll
                    for (int a = 0; a < NARR; a++) {
                        const double* arr = a == 0 ? in.pot : (a == 1 ? in.grad : in.hess);
                        o.r[h][a] = rsrc(arr + (size_t)lf.box * npbox * (a + 1) * nd + d);
                    }
                    // y phase of this lane's column (upward shift: conjugate of the table)
                    o.py[h] = bload_c(rP, oPy, lf.jy * rowb);
                    o.py[h].x *= wy * wl;
                    o.py[h].y *= -wy * wl;
                }
                return o;
            };
            struct YOp {
                cplx tt[NTAB];
                double a[LPC][NOUT];
            };
            auto yload = [&](const Pass& pa, int s) {
                YOp o;
#pragma unroll
                for (int u = 0; u < NTAB; u++) o.tt[u] = bload_c(rT, oTy, u * tbytes + 4 * s * rowb);
#pragma unroll
                for (int h = 0; h < LPC; h++)
#pragma unroll
                    for (int oo = 0; oo < NOUT; oo++)
                        o.a[h][oo] = bload_d(pa.r[h][stride_o[oo] - 1], oF[s] * stride_o[oo],
                                             comp_o[oo] * nd * (int)sizeof(double));
                return o;
            };
            if constexpr (NTAB == 1) {
                // one table: its rows stay in registers for the column, and
                // the leaf values of a whole pass are loaded one pass ahead
                // (the loop's loads are then the leaf values alone, waited
                // for in the order they were issued)
                cplx tt[NKS];
#pragma unroll
                for (int s = 0; s < NKS; s++) tt[s] = bload_c(rT, oTy, 4 * s * rowb);
                struct PV {
                    cplx py[LPC];
                    double a[NKS][LPC];
                };
                auto pv_at = [&](int li) {
                    const Pass pa = pass_at(li);
                    PV v;
#pragma unroll
                    for (int h = 0; h < LPC; h++) {
                        v.py[h] = pa.py[h];
#pragma unroll
                        for (int s = 0; s < NKS; s++) v.a[s][h] = bload_d(pa.r[h][0], oF[s], 0);
                    }
                    return v;
                };
                PV vc = pv_at(0);
                for (int li = 0; li < col.y; li += LPC) {
                    const PV vn = pv_at(min(li + LPC, col.y - 1));
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int s = 0; s < NKS; s++)
#pragma unroll
                        for (int h = 0; h < LPC; h++) {
                            const cplx ty = cmul(tt[s], vc.py[h]);
                            Gr[0] = mfma(vc.a[s][h], ty.x, Gr[0]);
                            Gi[0] = mfma(vc.a[s][h], ty.y, Gi[0]);
                        }
                    vc = vn;
                }
            } else {
            Pass pc = pass_at(0);
            YOp yc = yload(pc, 0);
            for (int li = 0; li < col.y; li += LPC) {
                // the table rows are read again by every pass (from L2)
                asm volatile("" ::: "memory");
                const Pass pn = pass_at(min(li + LPC, col.y - 1));
#pragma unroll
                for (int s = 0; s < NKS; s++) {
                    const YOp yn = s + 1 < NKS ? yload(pc, s + 1) : yload(pn, 0);
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int h = 0; h < LPC; h++) {
                        cplx ty[NTAB];
#pragma unroll
                        for (int u = 0; u < NTAB; u++) ty[u] = cmul(yc.tt[u], pc.py[h]);
#pragma unroll
                        for (int o = 0; o < NOUT; o++) {
                            Gr[chan_x(o)] = mfma(yc.a[h][o], ty[chan_y(o)].x, Gr[chan_x(o)]);
                            Gi[chan_x(o)] = mfma(yc.a[h][o], ty[chan_y(o)].y, Gi[chan_x(o)]);
                        }
                    }
                    yc = yn;
                }
                pc = pn;
            }
            }
#pragma unroll
            for (int t = 0; t < NTM; t++) {
                cplx px = bload_c(rP, oPx, (l0.jx * ps + 16 * t) * (int)sizeof(cplx));
                px.y = -px.y;
#pragma unroll
                for (int v = 0; v < NTAB; v++) {
#pragma unroll
                    for (int s = 0; s < NKS; s++) {
                        cplx tx = cmul(bload_c(rT, oTx, v * tbytes + 4 * s * rowb + 16 * t * (int)sizeof(cplx)), px);
                        acc[0][t] = mfma(tx.x, Gr[v][s], acc[0][t]);
                        acc[1][t] = mfma(tx.y, Gi[v][s], acc[1][t]);
                        acc[2][t] = mfma(tx.x, Gi[v][s], acc[2][t]);
                        acc[3][t] = mfma(tx.y, Gr[v][s], acc[3][t]);
                    }
                }
            }
        }
        if constexpr (BR) {
            // the waves that continue a piece hand their sums to its first wave
            if (valid && ch.w == 0) {
                double* mine = red + (wave - 1) * NACC * kWave;
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int t = 0; t < NTM; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) mine[((p * NTM + t) * 4 + r) * kWave + lane] = acc[p][t][r];
            }
            __syncthreads();
            for (int w = wave + 1; valid && w < wave + ch.w; w++) {
                const double* other = red + (w - 1) * NACC * kWave;
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int t = 0; t < NTM; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) acc[p][t][r] += other[((p * NTM + t) * 4 + r) * kWave + lane];
            }
            if constexpr (!DS) __syncthreads();  // the slots are refilled for the next density
        }
        if (!valid || ch.w == 0) continue;
        cplx* out = partial + ((size_t)ch.z * nd + d) * nexp;
#pragma unroll
        for (int t = 0; t < NTM; t++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int k1 = 16 * t + drow(q, r);
                if (k1 < Mlo && k2 < H) {
                    double AC = acc[0][t][r], BD = acc[1][t][r], AD = acc[2][t][r], BC = acc[3][t][r];
                    out[k1 * H + k2] = cplx{AC - BD, AD + BC};
                    int km = P - 1 - k1;
                    if (km > k1) out[km * H + k2] = cplx{AC + BD, AD - BC};
                }
            }
        }
    }
}

// Mirror-folded local expansion of a cutoff box: for x frequency m < Mh =
// ceil(npw/2) and retained y frequency k2 < H, S = L(m, k2) + L(npw-1-m, k2)
// and D = L(m, k2) - L(npw-1-m, k2) (S = L, D = 0 on the middle frequency of
// an odd expansion).  Stored [box][density][Mh4][H16] with
// Mh4 = 4 ceil(Mh/4), H16 = 16 ceil(H/16), zero in the padding.
struct Fold {
    cplx S, D;
};

__device__ __host__ inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Downward pass.  y frequencies are processed in tiles of 16:
//   FF_t^T = S^T a_t + i D^T b_t      a_t + i b_t = E_t diag(px), the x
//                                     contraction folded with
//                                     E_t(npw-1-m) = conj E_t(m)
//   out_o^T += Re[(E_y(o) diag(py w))^T FF_x(o)^T]
// for the NTAB-table channels o (chan_x, chan_y).  The second product reads
// the first one's accumulators directly.  E_t(m, i) is
// tab[level*lstride + (t*ps + m)*16 + i].  The forward evaluation
// (bfgt/bfgt_pwrouts.f:549-784) uses the PW-to-values tables and w = 1/2 on
// the middle y frequency of an odd expansion (:591-595); the adjoint of the
// forward transform is NTAB = 1 with the conjugated values-to-PW table and
// w = 1, writing the leaf-value cotangent.
//
// Downward pass over column groups: the PW leaves of one cutoff box, level
// and x cell jx share the x contraction FF (it depends only on the folded
// expansion, the level's tables and the x phase), so one wave forms FF once
// and then runs the y contraction and the output update for the group's
// leaves, LG at a time (the y-table loads of a step serve all LG leaves; a
// missing leaf is a zero-phase copy that is not stored).  groups[i] =
// (first leaf, count) over the leaf array sorted by (box, level, jx, jy).
// Per leaf the operands and products are those of the one-leaf formulation:
// FF_t^T = S^T a_t + i D^T b_t, out_o^T += Re[(E_y diag(py w))^T FF_x(o)^T].
// Both contractions load the operands of the next K-step before issuing the
// products of the current one, and all addresses are buffer offsets, so that
// the two-table kernel fits three waves per SIMD (165 registers).
template <int NTAB, int NTP, int LG>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_down_col(
    int n, int P, int ps, int nd, int ngroups, const int2* __restrict__ groups, const PwLeaf* __restrict__ leaves,
    const cplx* __restrict__ tab, int lstride, int ymid_half, const cplx* __restrict__ phase,
    const int* __restrict__ levbase, const Fold* __restrict__ fold, double* __restrict__ pot,
    double* __restrict__ grad, double* __restrict__ hess, double oscale, int accum) {
    constexpr int NOUT = nchan(NTAB);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= ngroups) return;  // no block-wide barriers in this kernel
    const int2 grp = groups[item];
    const int q = lane >> 4, cc = lane & 15;
    const int H = (P + 1) / 2, P2 = P / 2, npbox = n * n;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const int nks1 = Mh4 / 4, nT = H16 / 16;
    const int tbytes = 16 * ps * (int)sizeof(cplx);  // one table of a level
    const PwLeaf l0 = leaves[grp.x];
    const rsrc_t rE = rsrc(tab + (size_t)l0.level * lstride);  // [t][m][i]
    const rsrc_t rP = rsrc(phase + levbase[l0.level]);         // [cell][k]
    const int sPx = l0.jx * ps * (int)sizeof(cplx);
    const double w2 = 2.0 * oscale;
    // lane offsets (bytes): table row q, point cc; phase row q; folded
    // expansion row q, column cc of a y tile
    const unsigned oE = (unsigned)(q * 16 + cc) * sizeof(cplx);
    const unsigned oP = (unsigned)q * sizeof(cplx);
    const unsigned oF = (unsigned)(q * H16 + cc) * sizeof(Fold);
    for (int d = 0; d < nd; d++) {
        const rsrc_t rF = rsrc(fold + ((size_t)l0.anc * nd + d) * Mh4 * H16);
        for (int tg = 0; tg < nT; tg += NTP) {
            v4d FFr[NTP][NTAB], FFi[NTP][NTAB];
#pragma unroll
            for (int p = 0; p < NTP; p++)
#pragma unroll
                for (int t = 0; t < NTAB; t++) FFr[p][t] = FFi[p][t] = v4d{0, 0, 0, 0};
            // x contraction, one K-step (rows 4s..4s+3 of the folded
            // expansion) per iteration; the operands of step s + 1 are loaded
            // before the products of step s are issued (the scheduling
            // barrier keeps that order), so that their latency overlaps them
            struct XOp {
                cplx S[NTP], D[NTP];
                cplx e[NTAB];
                cplx px;
            };
            auto xload = [&](int s) {
                XOp o;
                const int sF = (4 * s * H16 + 16 * tg) * (int)sizeof(Fold);
#pragma unroll
                for (int p = 0; p < NTP; p++) {
                    o.S[p] = bload_c(rF, oF, sF + 16 * p * (int)sizeof(Fold));
                    o.D[p] = bload_c(rF, oF, sF + 16 * p * (int)sizeof(Fold) + (int)sizeof(cplx));
                }
#pragma unroll
                for (int t = 0; t < NTAB; t++) o.e[t] = bload_c(rE, oE, t * tbytes + s * 64 * (int)sizeof(cplx));
                o.px = bload_c(rP, oP, sPx + 4 * s * (int)sizeof(cplx));
                return o;
            };
            XOp xc = xload(0);
            for (int s = 0; s < nks1; s++) {
                const XOp xn = xload(min(s + 1, nks1 - 1));
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = 0; t < NTAB; t++) {
                    const cplx e = cmul(xc.e[t], xc.px);
#pragma unroll
                    for (int p = 0; p < NTP; p++) {
                        FFr[p][t] = mfma(xc.S[p].x, e.x, FFr[p][t]);
                        FFi[p][t] = mfma(xc.S[p].y, e.x, FFi[p][t]);
                        FFr[p][t] = mfma(-xc.D[p].y, e.y, FFr[p][t]);
                        FFi[p][t] = mfma(xc.D[p].x, e.y, FFi[p][t]);
                    }
                }
                xc = xn;
            }
            // y contraction: K-step k covers rows 16 tg + 4k + q of FF^T,
            // element k % 4 of FF[k / 4]; steps whose rows are all padding
            // are skipped (uniform).  Operands of step k + 1 (clamped to a
            // valid step) are loaded before the products of step k.
            const int nyk = min(4 * NTP, (H - 16 * tg + 3) / 4);
            for (int b = 0; b < grp.y; b += LG) {
                // the table rows are read again by every pass (from L2): kept
                // live across the passes they would hold 8 NTAB NTP registers
                asm volatile("" ::: "memory");
                int sPy[LG], box[LG];
                double wg[LG];
#pragma unroll
                for (int g = 0; g < LG; g++) {
                    const PwLeaf lf = leaves[grp.x + min(b + g, grp.y - 1)];
                    sPy[g] = (lf.jy * ps + 16 * tg) * (int)sizeof(cplx);
                    box[g] = lf.box;
                    wg[g] = b + g < grp.y ? 1.0 : 0.0;
                }
                struct YOp {
                    cplx ey[NTAB];
                    cplx py[LG];
                };
                auto yload = [&](int k) {
                    YOp o;
                    const int kk = min(k, nyk - 1);
#pragma unroll
                    for (int u = 0; u < NTAB; u++)
                        o.ey[u] = bload_c(rE, oE, u * tbytes + (16 * tg + 4 * kk) * 16 * (int)sizeof(cplx));
#pragma unroll
                    for (int g = 0; g < LG; g++) o.py[g] = bload_c(rP, oP, sPy[g] + 4 * kk * (int)sizeof(cplx));
                    return o;
                };
                v4d out[LG][NOUT];
#pragma unroll
                for (int g = 0; g < LG; g++)
#pragma unroll
                    for (int o = 0; o < NOUT; o++) out[g][o] = v4d{0, 0, 0, 0};
                YOp yc = yload(0);
#pragma unroll
                for (int k = 0; k < 4 * NTP; k++) {
                    YOp yn = yc;
                    if (k + 1 < 4 * NTP) yn = yload(k + 1);
                    __builtin_amdgcn_sched_barrier(0);
                    if (k < nyk) {
                        const int p = k / 4, s = k % 4;
                        const double w = (ymid_half && 16 * tg + 4 * k + q >= P2) ? 0.5 : 1.0;
#pragma unroll
                        for (int g = 0; g < LG; g++) {
                            cplx pyw = yc.py[g];
                            pyw.x *= w * wg[g];
                            pyw.y *= w * wg[g];
                            cplx eu[NTAB];
#pragma unroll
                            for (int u = 0; u < NTAB; u++) eu[u] = cmul(yc.ey[u], pyw);
#pragma unroll
                            for (int o = 0; o < NOUT; o++)
                                out[g][o] = mfma(eu[chan_y(o)].x, FFr[p][chan_x(o)][s], out[g][o]);
#pragma unroll
                            for (int o = 0; o < NOUT; o++)
                                out[g][o] = mfma(-eu[chan_y(o)].y, FFi[p][chan_x(o)][s], out[g][o]);
                        }
                    }
                    yc = yn;
                }
                // out^T holds [k2 = drow(q, r)][k1 = cc].  The outputs are
                // read for all (leaf, r) at clamped points first, so that one
                // memory round trip serves the whole batch, then added and
                // stored where valid.  A buffer resource per (leaf, channel),
                // the row in the lane offset.
                constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
                rsrc_t rO[LG][NOUT];
#pragma unroll
                for (int g = 0; g < LG; g++)
#pragma unroll
                    for (int o = 0; o < NOUT; o++) {
                        double* a = o == 0 ? pot : (o < 3 ? grad : hess);
                        rO[g][o] = rsrc(a + ((size_t)box[g] * npbox * stride_o[o] + comp_o[o]) * nd + d);
                    }
                unsigned orow[4];
#pragma unroll
                for (int r = 0; r < 4; r++)
                    orow[r] = (unsigned)(min(cc, n - 1) + n * min(drow(q, r), n - 1)) * (unsigned)(nd * sizeof(double));
                double old[LG][4][NOUT];
                if (accum) {
#pragma unroll
                    for (int g = 0; g < LG; g++)
#pragma unroll
                        for (int r = 0; r < 4; r++)
#pragma unroll
                            for (int o = 0; o < NOUT; o++) old[g][r][o] = bload_d(rO[g][o], orow[r] * stride_o[o], 0);
                } else {
#pragma unroll
                    for (int g = 0; g < LG; g++)
#pragma unroll
                        for (int r = 0; r < 4; r++)
#pragma unroll
                            for (int o = 0; o < NOUT; o++) old[g][r][o] = 0.0;
                }
                if (cc < n) {
#pragma unroll
                    for (int g = 0; g < LG; g++) {
                        if (b + g >= grp.y) break;
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            if (drow(q, r) < n) {
#pragma unroll
                                for (int o = 0; o < NOUT; o++)
                                    bstore_d(old[g][r][o] + w2 * out[g][o][r], rO[g][o], orow[r] * stride_o[o], 0);
                            }
                    }
                }
            }
        }
    }
}

// Downward pass for leaf orders n <= 8 over column pairs: the PW leaves of
// one cutoff box and level in x cells jxA and jxB fill the two 8-wide halves
// of the x contraction's point dimension, so FF is formed once per group;
// each y-contraction product then evaluates a 2 x 2 block of leaves (a
// quad): rows 0-7 of its point dimension are the y cell jyA, rows 8-15 jyB.
// qgroups[i] = (first quad, count, a leaf of column A, a leaf of column B or
// -1); quads[k] = leaves (A, jyA), (B, jyA), (A, jyB), (B, jyB) or -1.  A
// missing column or row gets zero phases; missing leaves are not stored.
// One wave per (group, density).
template <int NTAB, int NTP>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_down_quad(
    int n, int P, int ps, int nd, int ngroups, const int4* __restrict__ qgroups, const int4* __restrict__ quads,
    const PwLeaf* __restrict__ leaves, const cplx* __restrict__ tab, int lstride, int ymid_half,
    const cplx* __restrict__ phase, const int* __restrict__ levbase, const Fold* __restrict__ fold,
    double* __restrict__ pot, double* __restrict__ grad, double* __restrict__ hess, double oscale, int accum) {
    constexpr int NOUT = nchan(NTAB);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= ngroups * nd) return;  // no block-wide barriers in this kernel
    // one wave per (quad group, density): the densities of a group on one XCD
    const int4 g = qgroups[item / nd];
    const int d = item % nd;
    const int q = lane >> 4, cc = lane & 15;
    const int half = cc >> 3, ci = cc & 7;
    const int H = (P + 1) / 2, P2 = P / 2, npbox = n * n;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const int nks1 = Mh4 / 4, nT = H16 / 16;
    const int tstride = 16 * ps;
    const PwLeaf la = leaves[g.z];
    const int jxb = g.w >= 0 ? leaves[g.w].jx : la.jx;
    const cplx* E = tab + (size_t)la.level * lstride;  // [t][m][i]
    const cplx* ph = phase + levbase[la.level];
    // this lane's column: its x phase, zero in an absent column B
    const cplx* phx = ph + (half ? jxb : la.jx) * ps;
    const double wx = (half && g.w < 0) ? 0.0 : 1.0;
    const double w2 = 2.0 * oscale;
    constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
    double* dst[NOUT];
    dst[0] = pot;
    if constexpr (NTAB >= 2) dst[1] = dst[2] = grad;
    if constexpr (NTAB >= 3) dst[3] = dst[4] = dst[5] = hess;
    {  // this wave's density
        const Fold* F = fold + ((size_t)la.anc * nd + d) * Mh4 * H16;
        for (int tg = 0; tg < nT; tg += NTP) {
            v4d FFr[NTP][NTAB], FFi[NTP][NTAB];
#pragma unroll
            for (int p = 0; p < NTP; p++)
#pragma unroll
                for (int t = 0; t < NTAB; t++) FFr[p][t] = FFi[p][t] = v4d{0, 0, 0, 0};
            // two K-steps per pass (a step past the end reads the last rows
            // with weight zero)
            for (int s0 = 0; s0 < nks1; s0 += 2) {
#pragma unroll
                for (int s1 = 0; s1 < 2; s1++) {
                    const int s = min(s0 + s1, nks1 - 1);
                    const double ws = s0 + s1 < nks1 ? wx : 0.0;
                    const int m = 4 * s + q;
                    Fold sd[NTP];
#pragma unroll
                    for (int p = 0; p < NTP; p++) sd[p] = F[m * H16 + 16 * (tg + p) + cc];
                    cplx px = phx[m];
                    px.x *= ws;
                    px.y *= ws;
#pragma unroll
                    for (int t = 0; t < NTAB; t++) {
                        const cplx e = cmul(E[t * tstride + m * 16 + ci], px);
#pragma unroll
                        for (int p = 0; p < NTP; p++) {
                            FFr[p][t] = mfma(sd[p].S.x, e.x, FFr[p][t]);
                            FFi[p][t] = mfma(sd[p].S.y, e.x, FFi[p][t]);
                            FFr[p][t] = mfma(-sd[p].D.y, e.y, FFr[p][t]);
                            FFi[p][t] = mfma(sd[p].D.x, e.y, FFi[p][t]);
                        }
                    }
                }
            }
            for (int k = g.x; k < g.x + g.y; k++) {
                const int4 qd = quads[k];
                // this lane's row of the y contraction: y cell jyA (half 0) or jyB
                const int ra = qd.x >= 0 ? qd.x : qd.y, rb = qd.z >= 0 ? qd.z : qd.w;
                const int lrow = half ? (rb >= 0 ? rb : ra) : ra;
                const double wy = (half && rb < 0) ? 0.0 : 1.0;
                const cplx* phy = ph + leaves[lrow].jy * ps;
                v4d out[NOUT];
#pragma unroll
                for (int o = 0; o < NOUT; o++) out[o] = v4d{0, 0, 0, 0};
#pragma unroll
                for (int p = 0; p < NTP; p++) {
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        const int m2s = 16 * (tg + p) + 4 * s + q;
                        if (16 * (tg + p) + 4 * s >= H) break;  // rows of this step are all padding (uniform)
                        cplx pyw = phy[m2s];
                        const double w = ((ymid_half && m2s >= P2) ? 0.5 : 1.0) * wy;
                        pyw.x *= w;
                        pyw.y *= w;
                        cplx eu[NTAB];
#pragma unroll
                        for (int u = 0; u < NTAB; u++) eu[u] = cmul(E[u * tstride + m2s * 16 + ci], pyw);
#pragma unroll
                        for (int o = 0; o < NOUT; o++) out[o] = mfma(eu[chan_y(o)].x, FFr[p][chan_x(o)][s], out[o]);
#pragma unroll
                        for (int o = 0; o < NOUT; o++) out[o] = mfma(-eu[chan_y(o)].y, FFi[p][chan_x(o)][s], out[o]);
                    }
                }
                // element r of lane (q, cc) = D[row 4r + q][col cc]: leaf (column
                // cc >> 3, row (4r + q) >> 3 = r >> 1), point (ci, (4r + q) & 7)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int lid = (r >> 1) ? (half ? qd.w : qd.z) : (half ? qd.y : qd.x);
                    const int yp = drow(q, r) & 7;
                    if (lid >= 0 && ci < n && yp < n) {
                        const size_t pb = (size_t)leaves[lid].box * npbox + ci + (size_t)n * yp;
#pragma unroll
                        for (int o = 0; o < NOUT; o++) {
                            double& v = dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d];
                            v = (accum ? v : 0.0) + w2 * out[o][r];
                        }
                    }
                }
            }
        }
    }
}

// Downward pass for leaf orders n <= 8: one wave per PAIR of PW leaves with
// the same cutoff box and level (a missing partner is a zero-weight copy).
// The 16-wide point dimension holds leaf 0's points in lanes 0-7 and leaf 1's
// in lanes 8-15: the x contraction shares the folded expansion (same box)
// and the tables (same level) and differs only in the phase each lane's leaf
// applies; the y contraction's output tile holds the two leaves' diagonal
// 8 x 8 blocks (the off-diagonal blocks mix the leaves and are not stored).
template <int NTAB, int NTP>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_down_pair(
    int n, int P, int ps, int nd, int npair, const int2* __restrict__ pairs, const PwLeaf* __restrict__ leaves,
    const cplx* __restrict__ tab, int lstride, int ymid_half, const cplx* __restrict__ phase,
    const int* __restrict__ levbase, const Fold* __restrict__ fold, double* __restrict__ pot,
    double* __restrict__ grad, double* __restrict__ hess, double oscale, int accum) {
    constexpr int NOUT = nchan(NTAB);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= npair) return;  // no block-wide barriers in this kernel
    const int q = lane >> 4, cc = lane & 15;
    const int half = cc >> 3, ci = cc & 7;
    const int H = (P + 1) / 2, P2 = P / 2, npbox = n * n;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const int nks1 = Mh4 / 4, nT = H16 / 16;
    const int tstride = 16 * ps;
    const int2 pr = pairs[item];
    const bool two = pr.y >= 0;
    const PwLeaf lf0 = leaves[pr.x];
    const PwLeaf lfh = leaves[(half && two) ? pr.y : pr.x];  // this lane's leaf
    const double wh = (half && !two) ? 0.0 : 1.0;          // zero weight for a missing partner
    const cplx* E = tab + (size_t)lf0.level * lstride;     // [t][m][i], shared level
    const cplx* ph = phase + levbase[lf0.level];
    const cplx* phx = ph + lfh.jx * ps;
    const cplx* phy = ph + lfh.jy * ps;
    for (int d = 0; d < nd; d++) {
        const Fold* F = fold + ((size_t)lf0.anc * nd + d) * Mh4 * H16;
        v4d out[NOUT];
#pragma unroll
        for (int o = 0; o < NOUT; o++) out[o] = v4d{0, 0, 0, 0};
        for (int tg = 0; tg < nT; tg += NTP) {
            v4d FFr[NTP][NTAB], FFi[NTP][NTAB];
#pragma unroll
            for (int p = 0; p < NTP; p++)
#pragma unroll
                for (int t = 0; t < NTAB; t++) FFr[p][t] = FFi[p][t] = v4d{0, 0, 0, 0};
            for (int s = 0; s < nks1; s++) {
                const int m = 4 * s + q;
                Fold sd[NTP];
#pragma unroll
                for (int p = 0; p < NTP; p++) sd[p] = F[m * H16 + 16 * (tg + p) + cc];
                cplx px = phx[m];
                px.x *= wh;
                px.y *= wh;
#pragma unroll
                for (int t = 0; t < NTAB; t++) {
                    cplx e = cmul(E[t * tstride + m * 16 + ci], px);
#pragma unroll
                    for (int p = 0; p < NTP; p++) {
                        FFr[p][t] = mfma(sd[p].S.x, e.x, FFr[p][t]);
                        FFi[p][t] = mfma(sd[p].S.y, e.x, FFi[p][t]);
                        FFr[p][t] = mfma(-sd[p].D.y, e.y, FFr[p][t]);
                        FFi[p][t] = mfma(sd[p].D.x, e.y, FFi[p][t]);
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < NTP; p++) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int m2s = 16 * (tg + p) + 4 * s + q;
                    if (16 * (tg + p) + 4 * s >= H) break;  // rows of this step are all padding (uniform)
                    cplx pyw = phy[m2s];  // row k2' = cc: this lane's leaf
                    const double w = (ymid_half && m2s >= P2) ? 0.5 : 1.0;
                    pyw.x *= w;
                    pyw.y *= w;
                    cplx eu[NTAB];
#pragma unroll
                    for (int u = 0; u < NTAB; u++) eu[u] = cmul(E[u * tstride + m2s * 16 + ci], pyw);
#pragma unroll
                    for (int o = 0; o < NOUT; o++) out[o] = mfma(eu[chan_y(o)].x, FFr[p][chan_x(o)][s], out[o]);
#pragma unroll
                    for (int o = 0; o < NOUT; o++) out[o] = mfma(-eu[chan_y(o)].y, FFi[p][chan_x(o)][s], out[o]);
                }
            }
        }
        // out^T holds [k2' = drow(q, r)][k1' = cc]; a leaf's block has k2' and cc in the same half
        if (ci < n && (half == 0 || two)) {
            const PwLeaf& lfo = lfh;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k2p = drow(q, r);
                const int k2 = k2p & 7;
                if ((k2p >> 3) == half && k2 < n) {
                    size_t pb = (size_t)lfo.box * npbox + ci + (size_t)n * k2;
                    const double w2 = 2.0 * oscale;
                    auto put = [&](double& v, double x) { v = (accum ? v : 0.0) + w2 * x; };
                    put(pot[pb * nd + d], out[0][r]);
                    if constexpr (NTAB >= 2) {
                        put(grad[(pb * 2 + 0) * nd + d], out[1][r]);
                        put(grad[(pb * 2 + 1) * nd + d], out[2][r]);
                    }
                    if constexpr (NTAB >= 3) {
                        put(hess[(pb * 3 + 0) * nd + d], out[3][r]);
                        put(hess[(pb * 3 + 1) * nd + d], out[4][r]);
                        put(hess[(pb * 3 + 2) * nd + d], out[5][r]);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// split-FP32 variants of the leaf transforms (v_mfma_f32_16x16x4f32)
// ---------------------------------------------------------------------------
//
// Every FP64 operand x is split into x_hi = fl32(x), x_lo = fl32(x - x_hi); a
// product a*b becomes a_hi*b_hi + a_hi*b_lo + a_lo*b_hi in FP32 MFMAs that
// accumulate in FP32.  The table entries, phases (exp(i t x) in FP64) and
// their products are formed in FP64 before the split; a product whose second
// factor is an FP32 accumulator (the chained GEMM) uses two MFMAs.  Results
// are converted to FP64 where they leave the kernel: the PW expansions (the
// merge and translation stay FP64) and the pot/grad/hess accumulation.
//

// Forward upward pass, split-FP32: one wave per (chunk, y tile), with the
// column sharing of the FP64 pass -- the y contractions of a column's leaves
// are summed in the FP32 accumulator G, then one x contraction per column
// piece (cpc[chunk] = (first piece, count, slot), cols[piece] = (first
// leaf, count); the leaves of a piece share the level and the x cell).
template <int NKS, int NTM>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_up_split(
    int n, int P, int ps, int nd, int nds, int nTN, int nitems, const double* __restrict__ fvals,
    const PwLeaf* __restrict__ leaves, const int4* __restrict__ cpc, const int2* __restrict__ cols,
    const cplx* __restrict__ tab, int lstride, const cplx* __restrict__ phase, const int* __restrict__ levbase,
    cplx* __restrict__ partial) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= nitems) return;  // no block-wide barriers in this kernel
    // item = (chunk, density if nds > 1, y tile): with nds = nd one density
    // per wave, else the densities in turn
    const int tN = item % nTN, ci = item / nTN / nds;
    const int d0 = nds > 1 ? (item / nTN) % nds : 0, d1 = nds > 1 ? d0 + 1 : nd;
    const int q = lane >> 4, cc = lane & 15;
    const int H = (P + 1) / 2, Mlo = (P + 1) / 2, nexp = P * H, npbox = n * n;
    const int k2 = 16 * tN + cc;
    const int ccl = min(cc, n - 1);
    int pidx[NKS];
#pragma unroll
    for (int s = 0; s < NKS; s++) pidx[s] = ccl + n * min(4 * s + q, n - 1);
    const int4 ch = cpc[ci];
    for (int d = d0; d < d1; d++) {
        v4f acc[4][NTM];
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int t = 0; t < NTM; t++) acc[p][t] = v4f{0, 0, 0, 0};
        for (int c = 0; c < ch.y; c++) {
            const int2 col = cols[ch.x + c];
            const PwLeaf l0 = leaves[col.x];
            const cplx* T = tab + (size_t)l0.level * lstride;
            const cplx* ph = phase + levbase[l0.level];
            // G = sum over the column's leaves of F (T diag py): K index m2 = 4s + q
            v4f Gr = {0, 0, 0, 0}, Gi = {0, 0, 0, 0};
            for (int li = 0; li < col.y; li++) {
                const PwLeaf lf = leaves[col.x + li];
                const size_t pb0 = (size_t)lf.box * npbox;
                cplx py = ph[lf.jy * ps + k2];
                py.y = -py.y;
#pragma unroll
                for (int s = 0; s < NKS; s++) {
                    const int m2 = 4 * s + q;
                    cplx ty = cmul(T[m2 * ps + k2], py);
                    Split a = split(fvals[(pb0 + pidx[s]) * nd + d]);
                    Gr = mfma3(a, split(ty.x), Gr);
                    Gi = mfma3(a, split(ty.y), Gi);
                }
            }
            // out += (T diag px)^T G: K-step r has K index m1 = 4q + r
#pragma unroll
            for (int t = 0; t < NTM; t++) {
                const int k1 = 16 * t + cc;
                cplx px = ph[l0.jx * ps + k1];
                px.y = -px.y;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m1 = drow32(q, r);
                    cplx tx = cmul(T[m1 * ps + k1], px);
                    Split xr = split(tx.x), xi = split(tx.y);
                    acc[0][t] = mfma2(xr, Gr[r], acc[0][t]);
                    acc[1][t] = mfma2(xi, Gi[r], acc[1][t]);
                    acc[2][t] = mfma2(xr, Gi[r], acc[2][t]);
                    acc[3][t] = mfma2(xi, Gr[r], acc[3][t]);
                }
            }
        }
        cplx* out = partial + ((size_t)ch.z * nd + d) * nexp;
#pragma unroll
        for (int t = 0; t < NTM; t++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int k1 = 16 * t + drow32(q, r);
                if (k1 < Mlo && k2 < H) {
                    double AC = acc[0][t][r], BD = acc[1][t][r], AD = acc[2][t][r], BC = acc[3][t][r];
                    out[k1 * H + k2] = cplx{AC - BD, AD + BC};
                    int km = P - 1 - k1;
                    if (km > k1) out[km * H + k2] = cplx{AC + BD, AD - BC};
                }
            }
        }
    }
}

// Forward downward pass, split-FP32: one wave per column group (grp[item] =
// (first leaf, count), the leaves sharing the cutoff box, the level and the
// x cell), as in the FP64 pass -- the x contraction FF of every y tile is
// formed once (FP32 accumulators, NT tiles), then each leaf's y contraction
// and output update.
template <int NTAB, int NT>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_down_split(
    int n, int P, int ps, int nd, int ngrp, const int2* __restrict__ grp, const PwLeaf* __restrict__ leaves,
    const cplx* __restrict__ tab, int lstride, const cplx* __restrict__ phase, const int* __restrict__ levbase,
    const Fold* __restrict__ fold, double* __restrict__ pot, double* __restrict__ grad, double* __restrict__ hess,
    double oscale, int accum) {
    constexpr int NOUT = nchan(NTAB);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    // wave-uniform: schedule entries are then read with scalar loads
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= ngrp) return;  // no block-wide barriers in this kernel
    const int q = lane >> 4, cc = lane & 15;
    const int H = (P + 1) / 2, P2 = P / 2, npbox = n * n;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const int nks1 = Mh4 / 4;
    const int tstride = 16 * ps;
    const int2 g = grp[item];
    const PwLeaf l0 = leaves[g.x];
    const cplx* E = tab + (size_t)l0.level * lstride;
    const cplx* ph = phase + levbase[l0.level];
    const cplx* phx = ph + l0.jx * ps;
    constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
    double* dst[NOUT];
    dst[0] = pot;
    if constexpr (NTAB >= 2) dst[1] = dst[2] = grad;
    if constexpr (NTAB >= 3) dst[3] = dst[4] = dst[5] = hess;
    const double w2 = 2.0 * oscale;
    for (int d = 0; d < nd; d++) {
        const Fold* F = fold + ((size_t)l0.anc * nd + d) * Mh4 * H16;
        // x contraction of the group, every y tile
        v4f FFr[NTAB][NT], FFi[NTAB][NT];
#pragma unroll
        for (int t = 0; t < NTAB; t++)
#pragma unroll
            for (int tM = 0; tM < NT; tM++) FFr[t][tM] = FFi[t][tM] = v4f{0, 0, 0, 0};
        for (int s = 0; s < nks1; s++) {
            const int m = 4 * s + q;
            const cplx px = phx[m];
            Split er[NTAB], ei[NTAB];
#pragma unroll
            for (int t = 0; t < NTAB; t++) {
                cplx e = cmul(E[t * tstride + m * 16 + cc], px);
                er[t] = split(e.x);
                ei[t] = split(e.y);
            }
#pragma unroll
            for (int tM = 0; tM < NT; tM++) {
                const Fold sd = F[m * H16 + 16 * tM + cc];
                Split sr = split(sd.S.x), si = split(sd.S.y), dr = split(sd.D.x), di = neg(split(sd.D.y));
#pragma unroll
                for (int t = 0; t < NTAB; t++) {
                    FFr[t][tM] = mfma3(sr, er[t], FFr[t][tM]);
                    FFr[t][tM] = mfma3(di, ei[t], FFr[t][tM]);
                    FFi[t][tM] = mfma3(si, er[t], FFi[t][tM]);
                    FFi[t][tM] = mfma3(dr, ei[t], FFi[t][tM]);
                }
            }
        }
        for (int li = 0; li < g.y; li++) {
            const PwLeaf lf = leaves[g.x + li];
            const cplx* phy = ph + lf.jy * ps;
            v4f out[NOUT];
#pragma unroll
            for (int o = 0; o < NOUT; o++) out[o] = v4f{0, 0, 0, 0};
#pragma unroll
            for (int tM = 0; tM < NT; tM++) {
                // register r of FF holds FF^T[16 tM + 4q + r][cc]: K-step r, K index 4q + r
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m2s = 16 * tM + drow32(q, r);
                    cplx pyw = phy[m2s];
                    const double w = (m2s >= P2) ? 0.5 : 1.0;
                    pyw.x *= w;
                    pyw.y *= w;
                    Split eur[NTAB], eui[NTAB];
#pragma unroll
                    for (int u = 0; u < NTAB; u++) {
                        cplx eu = cmul(E[u * tstride + m2s * 16 + cc], pyw);
                        eur[u] = split(eu.x);
                        eui[u] = split(-eu.y);
                    }
#pragma unroll
                    for (int o = 0; o < NOUT; o++) {
                        out[o] = mfma2(eur[chan_y(o)], FFr[chan_x(o)][tM][r], out[o]);
                        out[o] = mfma2(eui[chan_y(o)], FFi[chan_x(o)][tM][r], out[o]);
                    }
                }
            }
            // register r of out^T holds [k2 = 4q + r][k1 = cc].  The outputs
            // are read at clamped points for all r first (one memory round
            // trip), then added and stored where valid.
            double old[4][NOUT];
            if (accum) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const size_t pb = (size_t)lf.box * npbox + min(cc, n - 1) + (size_t)n * min(drow32(q, r), n - 1);
#pragma unroll
                    for (int o = 0; o < NOUT; o++) old[r][o] = dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int o = 0; o < NOUT; o++) old[r][o] = 0.0;
            }
            if (cc < n) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int k2 = drow32(q, r);
                    if (k2 < n) {
                        const size_t pb = (size_t)lf.box * npbox + cc + (size_t)n * k2;
#pragma unroll
                        for (int o = 0; o < NOUT; o++)
                            dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d] = old[r][o] + w2 * (double)out[o][r];
                    }
                }
            }
        }
    }
}

// Split-FP32 counterparts of k_up_quad and k_down_quad (n <= 8): the same
// 2 x 2 leaf blocks over column pairs.  The FP32 accumulator holds D[4q +
// r][c] (the FP64 one D[4r + q][c]), so a product that reads an accumulator
// as its B operand takes K index 4q + r in K-step r, and the A operand is
// loaded at that index.
template <int NTM>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_up_split_quad(
    int n, int P, int ps, int nd, int nds, int nTN, int nitems, const double* __restrict__ fvals,
    const PwLeaf* __restrict__ leaves, const int4* __restrict__ chunks, const int4* __restrict__ segs,
    const int4* __restrict__ quads, const cplx* __restrict__ tab, int lstride, const cplx* __restrict__ phase,
    const int* __restrict__ levbase, cplx* __restrict__ partial) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= nitems) return;  // no block-wide barriers in this kernel
    const int tN = item % nTN, ci = item / nTN / nds;  // as in k_up_split
    const int d0 = nds > 1 ? (item / nTN) % nds : 0, d1 = nds > 1 ? d0 + 1 : nd;
    const int q = lane >> 4, cc = lane & 15;
    const int half = cc >> 3, xi = min(cc & 7, n - 1);
    const int H = (P + 1) / 2, Mlo = (P + 1) / 2, nexp = P * H, npbox = n * n;
    const int k2 = 16 * tN + cc;
    const int4 ch = chunks[ci];
    for (int d = d0; d < d1; d++) {
        v4f acc[4][NTM];
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int t = 0; t < NTM; t++) acc[p][t] = v4f{0, 0, 0, 0};
        for (int sg = ch.x; sg < ch.x + ch.y; sg++) {
            const int4 seg = segs[sg];
            const PwLeaf la = leaves[seg.z];
            const int jxb = seg.w >= 0 ? leaves[seg.w].jx : la.jx;
            const double wb = seg.w >= 0 ? 1.0 : 0.0;
            const cplx* T = tab + (size_t)la.level * lstride;
            const cplx* ph = phase + levbase[la.level];
            // G = sum over the quads of F (T diag py): rows (column, x point),
            // K (y cell, y point), one product for four leaves
            v4f Gr = v4f{0, 0, 0, 0}, Gi = v4f{0, 0, 0, 0};
            for (int k = seg.x; k < seg.x + seg.y; k++) {
                const int4 qd = quads[k];
                const int ra = qd.x >= 0 ? qd.x : qd.y, rb = qd.z >= 0 ? qd.z : qd.w;
                const int jya = leaves[ra].jy, jyb = rb >= 0 ? leaves[rb].jy : jya;
                const int lfa = half ? qd.y : qd.x, lfb = half ? qd.w : qd.z;
                const size_t pba = (size_t)leaves[lfa >= 0 ? lfa : ra].box * npbox;
                const size_t pbb = (size_t)leaves[lfb >= 0 ? lfb : ra].box * npbox;
                const double wa = lfa >= 0 ? 1.0 : 0.0, wbr = lfb >= 0 ? 1.0 : 0.0;
                cplx pya = ph[jya * ps + k2], pyb = ph[jyb * ps + k2];
                pya.y = -pya.y;  // upward shift: conjugate of the table
                pyb.y = -pyb.y;
                if (rb < 0) pyb = cplx{0.0, 0.0};
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int yp = 4 * (s & 1) + q;  // y point of this K index
                    const cplx py = (s >> 1) ? pyb : pya;
                    const cplx ty = cmul(T[min(yp, 15) * ps + k2], py);
                    const Split a = split(((s >> 1) ? wbr : wa) *
                                          fvals[(((s >> 1) ? pbb : pba) + xi + (size_t)n * min(yp, n - 1)) * nd + d]);
                    Gr = mfma3(a, split(ty.x), Gr);
                    Gi = mfma3(a, split(ty.y), Gi);
                }
            }
            // out += (T diag px)^T G: K-step r reads G's element r, K index
            // 4q + r = (column q >> 1, x point 4 (q & 1) + r)
#pragma unroll
            for (int t = 0; t < NTM; t++) {
                const int k1 = 16 * t + cc;
                cplx pxa = ph[la.jx * ps + k1], pxb = ph[jxb * ps + k1];
                pxa.y = -pxa.y;
                pxb.x *= wb;
                pxb.y *= -wb;
                const cplx px = (q >> 1) ? pxb : pxa;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int xp = 4 * (q & 1) + r;
                    const cplx tx = cmul(T[xp * ps + k1], px);
                    const Split xr = split(tx.x), xi2 = split(tx.y);
                    acc[0][t] = mfma2(xr, Gr[r], acc[0][t]);
                    acc[1][t] = mfma2(xi2, Gi[r], acc[1][t]);
                    acc[2][t] = mfma2(xr, Gi[r], acc[2][t]);
                    acc[3][t] = mfma2(xi2, Gr[r], acc[3][t]);
                }
            }
        }
        cplx* out = partial + ((size_t)ch.z * nd + d) * nexp;
#pragma unroll
        for (int t = 0; t < NTM; t++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int k1 = 16 * t + drow32(q, r);
                if (k1 < Mlo && k2 < H) {
                    double AC = acc[0][t][r], BD = acc[1][t][r], AD = acc[2][t][r], BC = acc[3][t][r];
                    out[k1 * H + k2] = cplx{AC - BD, AD + BC};
                    int km = P - 1 - k1;
                    if (km > k1) out[km * H + k2] = cplx{AC + BD, AD - BC};
                }
            }
        }
    }
}

template <int NTAB, int NT>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_down_split_quad(
    int n, int P, int ps, int nd, int ngroups, const int4* __restrict__ qgroups, const int4* __restrict__ quads,
    const PwLeaf* __restrict__ leaves, const cplx* __restrict__ tab, int lstride, const cplx* __restrict__ phase,
    const int* __restrict__ levbase, const Fold* __restrict__ fold, double* __restrict__ pot,
    double* __restrict__ grad, double* __restrict__ hess, double oscale, int accum) {
    constexpr int NOUT = nchan(NTAB);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= ngroups * nd) return;  // no block-wide barriers in this kernel
    const int4 g = qgroups[item / nd];
    const int d = item % nd;
    const int q = lane >> 4, cc = lane & 15;
    const int half = cc >> 3, ci = cc & 7;
    const int H = (P + 1) / 2, P2 = P / 2, npbox = n * n;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const int nks1 = Mh4 / 4;
    const int tstride = 16 * ps;
    const PwLeaf la = leaves[g.z];
    const int jxb = g.w >= 0 ? leaves[g.w].jx : la.jx;
    const cplx* E = tab + (size_t)la.level * lstride;
    const cplx* ph = phase + levbase[la.level];
    const cplx* phx = ph + (half ? jxb : la.jx) * ps;
    const double wx = (half && g.w < 0) ? 0.0 : 1.0;
    const double w2 = 2.0 * oscale;
    constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
    double* dst[NOUT];
    dst[0] = pot;
    if constexpr (NTAB >= 2) dst[1] = dst[2] = grad;
    if constexpr (NTAB >= 3) dst[3] = dst[4] = dst[5] = hess;
    const Fold* F = fold + ((size_t)la.anc * nd + d) * Mh4 * H16;
    // x contraction of both columns, every y tile
    v4f FFr[NTAB][NT], FFi[NTAB][NT];
#pragma unroll
    for (int t = 0; t < NTAB; t++)
#pragma unroll
        for (int tM = 0; tM < NT; tM++) FFr[t][tM] = FFi[t][tM] = v4f{0, 0, 0, 0};
    for (int s = 0; s < nks1; s++) {
        const int m = 4 * s + q;
        cplx px = phx[m];
        px.x *= wx;
        px.y *= wx;
        Split er[NTAB], ei[NTAB];
#pragma unroll
        for (int t = 0; t < NTAB; t++) {
            const cplx e = cmul(E[t * tstride + m * 16 + ci], px);
            er[t] = split(e.x);
            ei[t] = split(e.y);
        }
#pragma unroll
        for (int tM = 0; tM < NT; tM++) {
            const Fold sd = F[m * H16 + 16 * tM + cc];
            Split sr = split(sd.S.x), si = split(sd.S.y), dr = split(sd.D.x), di = neg(split(sd.D.y));
#pragma unroll
            for (int t = 0; t < NTAB; t++) {
                FFr[t][tM] = mfma3(sr, er[t], FFr[t][tM]);
                FFr[t][tM] = mfma3(di, ei[t], FFr[t][tM]);
                FFi[t][tM] = mfma3(si, er[t], FFi[t][tM]);
                FFi[t][tM] = mfma3(dr, ei[t], FFi[t][tM]);
            }
        }
    }
    for (int k = g.x; k < g.x + g.y; k++) {
        const int4 qd = quads[k];
        // this lane's row of the y contraction: y cell jyA (half 0) or jyB
        const int ra = qd.x >= 0 ? qd.x : qd.y, rb = qd.z >= 0 ? qd.z : qd.w;
        const int lrow = half ? (rb >= 0 ? rb : ra) : ra;
        const double wy = (half && rb < 0) ? 0.0 : 1.0;
        const cplx* phy = ph + leaves[lrow].jy * ps;
        v4f out[NOUT];
#pragma unroll
        for (int o = 0; o < NOUT; o++) out[o] = v4f{0, 0, 0, 0};
#pragma unroll
        for (int tM = 0; tM < NT; tM++) {
            // register r of FF holds FF^T[16 tM + 4q + r][cc]: K-step r, K index 4q + r
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m2s = 16 * tM + drow32(q, r);
                cplx pyw = phy[m2s];
                const double w = (m2s >= P2 ? 0.5 : 1.0) * wy;
                pyw.x *= w;
                pyw.y *= w;
                Split eur[NTAB], eui[NTAB];
#pragma unroll
                for (int u = 0; u < NTAB; u++) {
                    const cplx eu = cmul(E[u * tstride + m2s * 16 + ci], pyw);
                    eur[u] = split(eu.x);
                    eui[u] = split(-eu.y);
                }
#pragma unroll
                for (int o = 0; o < NOUT; o++) {
                    out[o] = mfma2(eur[chan_y(o)], FFr[chan_x(o)][tM][r], out[o]);
                    out[o] = mfma2(eui[chan_y(o)], FFi[chan_x(o)][tM][r], out[o]);
                }
            }
        }
        // element r of lane (q, cc) = D[row 4q + r][col cc]: leaf (column
        // cc >> 3, row q >> 1), point (ci, 4 (q & 1) + r)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int lid = (q >> 1) ? (half ? qd.w : qd.z) : (half ? qd.y : qd.x);
            const int yp = drow32(q, r) & 7;
            if (lid >= 0 && ci < n && yp < n) {
                const size_t pb = (size_t)leaves[lid].box * npbox + ci + (size_t)n * yp;
#pragma unroll
                for (int o = 0; o < NOUT; o++) {
                    double& v = dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d];
                    v = (accum ? v : 0.0) + w2 * (double)out[o][r];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// three-slice bf16 variants of the leaf transforms (v_mfma_f32_16x16x16_bf16)
// ---------------------------------------------------------------------------
//
// BFGT_PREC_BF16X3: every operand is cut into three bf16 slices and each
// product formed from six slice products accumulated in FP32 (mfma6).  The
// point and frequency contractions are 16 long, one K block of the bf16
// instruction; its accumulator's element r of lane (q, c) is row 4q + r, the
// four K values lane (q, c) holds as the B operand of the next product, so
// the two contractions of each pass chain through registers as in the FP64
// kernels.  Schedules as the split-FP32 column kernels (n <= 16, one wave per
// (chunk, y tile) upward, per column group downward); tables, phases and
// their products in FP64, the merge, translation and output accumulation in
// FP64.

// The phased tables of both passes depend only on the level, the cell of
// the leaf inside its cutoff box and the frequencies, so they are formed and
// sliced once per plan (k_b3_tables), per cell: the cells of level l are
// numbered (2^(l - L0) - 1) + j, j the cell's x or y index.  Each entry is
// the six operand registers (slices h, m, l of the real part, then of the
// second part) of one lane of one 16 x 16 block, stored lane-contiguous:
//   kind 0, down x (B):  E_t[m][i] px_j[m]         (K m = 16 b + 4q + k, col i = c)
//   kind 1, down y (A):  E_u[k2][i] py_j[k2] w(k2)  (row i = c, K k2 = 16 b + 4q + k; real, -imag)
//   kind 2, up y (B):    T[m2][k2] conj(py_j[k2])   (K m2 = 4q + k, col k2 = 16 b + c)
//   kind 3, up x (A):    T[m1][k1] conj(px_j[k1])   (row k1 = 16 b + c, K m1 = 4q + k)
// zero past the retained frequencies (H) where those meet nonzero operands.
__device__ inline int b3_cell(int level, int lev0, int j) { return (1 << (level - lev0)) - 1 + j; }
__device__ inline void b3_load(const v4s* __restrict__ tab, size_t entry, int lane, Slices3& a, Slices3& b) {
    const v4s* e = tab + entry * 6 * 64 + lane;
    a.h = e[0];
    a.m = e[64];
    a.l = e[128];
    b.h = e[192];
    b.m = e[256];
    b.l = e[320];
}

__global__ void k_b3_tables(int kind, int P, int ps, int lev0, int ncell, int ntab, int nblk,
                            const cplx* __restrict__ tup_f, const cplx* __restrict__ tdn_f,
                            const cplx* __restrict__ phase, const int* __restrict__ levbase, v4s* __restrict__ out) {
    const int H = (P + 1) / 2, P2 = P / 2;
    const long tot = (long)ncell * ntab * nblk * 64;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int lane = (int)(e % 64);
        long r = e / 64;
        const int b = (int)(r % nblk);
        r /= nblk;
        const int t = (int)(r % ntab);
        const int cell = (int)(r / ntab);
        const int lo = 31 - __clz(cell + 1);
        const int l = lev0 + lo, j = cell + 1 - (1 << lo);
        const cplx* ph = phase + levbase[l] + (size_t)j * ps;
        const cplx* E = tdn_f + (size_t)l * 3 * 16 * ps + (size_t)t * 16 * ps;  // [m][i]
        const cplx* T = tup_f + (size_t)l * 16 * ps;                             // [m][k]
        const int q = lane >> 4, c = lane & 15;
        double re[4], im[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            cplx v{0.0, 0.0};
            if (kind == 0) {
                const int m = 16 * b + 4 * q + k;
                if (m < H) v = cmul(E[m * 16 + c], ph[m]);
            } else if (kind == 1) {
                const int k2 = 16 * b + 4 * q + k;
                if (k2 < H) {
                    v = cmul(E[k2 * 16 + c], ph[k2]);
                    const double w = k2 >= P2 ? 0.5 : 1.0;
                    v = cplx{w * v.x, -w * v.y};
                }
            } else if (kind == 2) {
                const int k2 = 16 * b + c;
                const cplx p = ph[k2];
                v = cmul(T[(4 * q + k) * ps + k2], cplx{p.x, -p.y});
            } else {
                const int k1 = 16 * b + c;
                const cplx p = ph[k1];
                v = cmul(T[(4 * q + k) * ps + k1], cplx{p.x, -p.y});
            }
            re[k] = v.x;
            im[k] = v.y;
        }
        const Slices3 sr = slice3(re), si = slice3(im);
        v4s* o = out + (size_t)(r * nblk + b) * 6 * 64 + lane;
        o[0] = sr.h;
        o[64] = sr.m;
        o[128] = sr.l;
        o[192] = si.h;
        o[256] = si.m;
        o[320] = si.l;
    }
}

// Upward pass: per column piece G = sum over its leaves of F (T diag py)
// (rows: x points, K: y points; the leaf values sliced per leaf, the phased
// table slices read per leaf cell), then acc += (T diag px)^T G (K: x
// points; G sliced once per column, the phased x table read per column).
template <int NTM, bool BR>
__global__ __launch_bounds__(kWave * kUpReduceWaves) void k_up_b3(
    int n, int P, int nd, int nds, int nTN, int nchunks, const double* __restrict__ fvals,
    const PwLeaf* __restrict__ leaves, const int4* __restrict__ cpc, const int2* __restrict__ cols, int lev0,
    const v4s* __restrict__ uy, const v4s* __restrict__ ux, cplx* __restrict__ partial) {
    constexpr int NACC = 16 * NTM;  // accumulator floats per lane
    // with BR the sums of waves 1.. of a workgroup, as in k_up_col
    __shared__ float red[BR ? (kUpReduceWaves - 1) * NACC * kWave : 1];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    // workgroup b: y tile b % nTN (and density (b / nTN) % nd with nds = nd)
    // of kUpReduceWaves consecutive chunks, XCD-contiguous as in k_up_col
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int tN = lb % nTN;
    const int d0 = nds > 1 ? (lb / nTN) % nds : 0, d1 = nds > 1 ? d0 + 1 : nd;
    const int ci = __builtin_amdgcn_readfirstlane((lb / nTN / nds) * kUpReduceWaves + wave);
    const bool valid = ci < nchunks;  // waves past the end still meet the barriers
    const int q = lane >> 4, cc = lane & 15;
    const int H = (P + 1) / 2, Mlo = (P + 1) / 2, nexp = P * H, npbox = n * n;
    const int k2 = 16 * tN + cc;
    // this lane's leaf points: x = cc, y = 4q + k (clamped; rows y >= n and
    // columns x >= n meet zero table rows)
    int pidx[4];
#pragma unroll
    for (int k = 0; k < 4; k++) pidx[k] = min(cc, n - 1) + n * min(4 * q + k, n - 1);
    const int4 ch = cpc[valid ? ci : 0];
    for (int d = d0; d < d1; d++) {
        v4f acc[4][NTM];
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int t = 0; t < NTM; t++) acc[p][t] = v4f{0, 0, 0, 0};
        for (int c = 0; valid && c < ch.y; c++) {
            const int2 col = cols[ch.x + c];
            const PwLeaf l0 = leaves[col.x];
            v4f Gr = {0, 0, 0, 0}, Gi = {0, 0, 0, 0};
            // a leaf's operands: its phased y table slices and values
            struct UOp {
                Slices3 br, bi;
                double a[4];
            };
            auto uop_at = [&](int li) {
                const PwLeaf lf = leaves[col.x + min(li, col.y - 1)];
                UOp o;
                b3_load(uy, (size_t)b3_cell(lf.level, lev0, lf.jy) * nTN + tN, lane, o.br, o.bi);
                const size_t pb0 = (size_t)lf.box * npbox;
#pragma unroll
                for (int k = 0; k < 4; k++) o.a[k] = fvals[(pb0 + pidx[k]) * nd + d];
                return o;
            };
            // the next leaf's operands are loaded into the same registers
            // once this leaf's products are issued
            UOp uc = uop_at(0);
            for (int li = 0; li < col.y; li++) {
                const Slices3 sa = slice3(uc.a);
                Gr = mfma6(sa, uc.br, Gr);
                Gi = mfma6(sa, uc.bi, Gi);
                __builtin_amdgcn_sched_barrier(0);
                if (li + 1 < col.y) uc = uop_at(li + 1);
                __builtin_amdgcn_sched_barrier(0);
            }
            // G's element k on lane (q, c) = G[x point 4q + k][k2]: the B operand as it stands
            const Slices3 gr = slice3(Gr[0], Gr[1], Gr[2], Gr[3]), gi = slice3(Gi[0], Gi[1], Gi[2], Gi[3]);
            const size_t ex = (size_t)b3_cell(l0.level, lev0, l0.jx) * NTM;
#pragma unroll
            for (int t = 0; t < NTM; t++) {
                Slices3 sxr, sxi;
                b3_load(ux, ex + t, lane, sxr, sxi);
                acc[0][t] = mfma6(sxr, gr, acc[0][t]);  // AC
                acc[1][t] = mfma6(sxi, gi, acc[1][t]);  // BD
                acc[2][t] = mfma6(sxr, gi, acc[2][t]);  // AD
                acc[3][t] = mfma6(sxi, gr, acc[3][t]);  // BC
            }
        }
        if constexpr (BR) {
            // the waves that continue a piece hand their sums to its first wave
            if (valid && ch.w == 0) {
                float* mine = red + (wave - 1) * NACC * kWave;
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int t = 0; t < NTM; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) mine[((p * NTM + t) * 4 + r) * kWave + lane] = acc[p][t][r];
            }
            __syncthreads();
            for (int w = wave + 1; valid && w < wave + ch.w; w++) {
                const float* other = red + (w - 1) * NACC * kWave;
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int t = 0; t < NTM; t++)
#pragma unroll
                        for (int r = 0; r < 4; r++) acc[p][t][r] += other[((p * NTM + t) * 4 + r) * kWave + lane];
            }
            __syncthreads();  // the slots are refilled for the next density
        }
        if (!valid || ch.w == 0) continue;
        cplx* out = partial + ((size_t)ch.z * nd + d) * nexp;
#pragma unroll
        for (int t = 0; t < NTM; t++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int k1 = 16 * t + drow32(q, r);
                if (k1 < Mlo && k2 < H) {
                    double AC = acc[0][t][r], BD = acc[1][t][r], AD = acc[2][t][r], BC = acc[3][t][r];
                    out[k1 * H + k2] = cplx{AC - BD, AD + BC};
                    int km = P - 1 - k1;
                    if (km > k1) out[km * H + k2] = cplx{AC + BD, AD - BC};
                }
            }
        }
    }
}

// Downward pass: per column group FF_t = S^T a_t + i D^T b_t (rows: y
// frequencies of a tile, K: x frequencies in blocks of 16; the folded
// expansion sliced per group, the phased x tables read per group), then per
// leaf out_o += Re[(E_y(o) diag(py w))^T FF_x(o)] (K: the tile's y
// frequencies; FF sliced once per group, the phased y tables read per leaf).
template <int NTAB, int NT>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_down_b3(
    int n, int P, int nd, int ngrp, const int2* __restrict__ grp, const PwLeaf* __restrict__ leaves, int lev0,
    int nkb, const v4s* __restrict__ dx, const v4s* __restrict__ dy, const Fold* __restrict__ fold,
    double* __restrict__ pot, double* __restrict__ grad, double* __restrict__ hess, double oscale, int accum) {
    constexpr int NOUT = nchan(NTAB);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int item = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave);
    if (item >= ngrp) return;  // no block-wide barriers in this kernel
    const int q = lane >> 4, cc = lane & 15;
    const int H = (P + 1) / 2, npbox = n * n;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const int2 g = grp[item];
    const PwLeaf l0 = leaves[g.x];
    const size_t ex = (size_t)b3_cell(l0.level, lev0, l0.jx) * NTAB;
    constexpr int stride_o[6] = {1, 2, 2, 3, 3, 3}, comp_o[6] = {0, 0, 1, 0, 1, 2};
    double* dst[NOUT];
    dst[0] = pot;
    if constexpr (NTAB >= 2) dst[1] = dst[2] = grad;
    if constexpr (NTAB >= 3) dst[3] = dst[4] = dst[5] = hess;
    const double w2 = 2.0 * oscale;
    for (int d = 0; d < nd; d++) {
        const Fold* F = fold + ((size_t)l0.anc * nd + d) * Mh4 * H16;
        v4f FFr[NTAB][NT], FFi[NTAB][NT];
#pragma unroll
        for (int t = 0; t < NTAB; t++)
#pragma unroll
            for (int tM = 0; tM < NT; tM++) FFr[t][tM] = FFi[t][tM] = v4f{0, 0, 0, 0};
        for (int kb = 0; kb < nkb; kb++) {
            // this lane's x frequencies m = 16 kb + 4q + k: rows past the
            // folded expansion read its last row, against zero table rows
            int mk[4];
#pragma unroll
            for (int k = 0; k < 4; k++) mk[k] = min(16 * kb + 4 * q + k, Mh4 - 1);
            Slices3 er[NTAB], ei[NTAB];
#pragma unroll
            for (int t = 0; t < NTAB; t++) b3_load(dx, (ex + t) * nkb + kb, lane, er[t], ei[t]);
#pragma unroll
            for (int tM = 0; tM < NT; tM++) {
                double sr[4], si[4], dr[4], di[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const Fold f = F[mk[k] * H16 + 16 * tM + cc];
                    sr[k] = f.S.x;
                    si[k] = f.S.y;
                    dr[k] = f.D.x;
                    di[k] = -f.D.y;
                }
                const Slices3 ssr = slice3(sr), ssi = slice3(si), sdr = slice3(dr), sdi = slice3(di);
#pragma unroll
                for (int t = 0; t < NTAB; t++) {
                    FFr[t][tM] = mfma6(ssr, er[t], FFr[t][tM]);
                    FFr[t][tM] = mfma6(sdi, ei[t], FFr[t][tM]);
                    FFi[t][tM] = mfma6(ssi, er[t], FFi[t][tM]);
                    FFi[t][tM] = mfma6(sdr, ei[t], FFi[t][tM]);
                }
            }
        }
        // FF's element k on lane (q, c) = FF[y frequency 16 tM + 4q + k][x point c]
        Slices3 bFr[NTAB][NT], bFi[NTAB][NT];
#pragma unroll
        for (int t = 0; t < NTAB; t++)
#pragma unroll
            for (int tM = 0; tM < NT; tM++) {
                bFr[t][tM] = slice3(FFr[t][tM][0], FFr[t][tM][1], FFr[t][tM][2], FFr[t][tM][3]);
                bFi[t][tM] = slice3(FFi[t][tM][0], FFi[t][tM][1], FFi[t][tM][2], FFi[t][tM][3]);
            }
        for (int li = 0; li < g.y; li++) {
            const PwLeaf lf = leaves[g.x + li];
            const size_t ey = (size_t)b3_cell(lf.level, lev0, lf.jy) * NTAB;
            v4f out[NOUT];
#pragma unroll
            for (int o = 0; o < NOUT; o++) out[o] = v4f{0, 0, 0, 0};
#pragma unroll
            for (int tM = 0; tM < NT; tM++) {
                Slices3 ar[NTAB], ai[NTAB];
#pragma unroll
                for (int u = 0; u < NTAB; u++) b3_load(dy, (ey + u) * NT + tM, lane, ar[u], ai[u]);
#pragma unroll
                for (int o = 0; o < NOUT; o++) {
                    out[o] = mfma6(ar[chan_y(o)], bFr[chan_x(o)][tM], out[o]);
                    out[o] = mfma6(ai[chan_y(o)], bFi[chan_x(o)][tM], out[o]);
                }
            }
            // element r of out on lane (q, c): y point 4q + r, x point c
            if (cc < n) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int yp = drow32(q, r);
                    if (yp < n) {
                        const size_t pb = (size_t)lf.box * npbox + cc + (size_t)n * yp;
#pragma unroll
                        for (int o = 0; o < NOUT; o++) {
                            double& v = dst[o][(pb * stride_o[o] + comp_o[o]) * nd + d];
                            v = (accum ? v : 0.0) + w2 * (double)out[o][r];
                        }
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// vector-FMA kernels (leaf orders above 16 or expansions above 128 waves)
// ---------------------------------------------------------------------------

constexpr int kVecThreads = 256;

template <int kUpMaxE>
__global__ __launch_bounds__(kVecThreads) void k_up_vec(int n, int npw, int nd, const double* __restrict__ fvals,
                                                        const PwLeaf* __restrict__ leaves,
                                                        const UpChunk* __restrict__ chunks,
                                                        const cplx* __restrict__ tpoly, const double* __restrict__ ts,
                                                        cplx* __restrict__ partial) {
    extern __shared__ double smem[];
    const int H = (npw + 1) / 2;
    const int nexp = npw * H;
    const int npbox = n * n;
    double* sf = smem;                               // n*n
    cplx* sT = reinterpret_cast<cplx*>(sf + npbox);  // n*npw : sT[m*npw + k]
    cplx* sff = sT + npw * n;                        // n*H : sff[m1 + n*k2]
    cplx* spx = sff + n * H;                         // npw
    cplx* spy = spx + npw;                           // H
    const UpChunk ch = chunks[blockIdx.x];
    const int tid = threadIdx.x;
    // blockIdx.y selects the segment of kUpMaxE * kVecThreads expansion
    // entries this block accumulates (long expansions take several)
    const int e0 = blockIdx.y * kUpMaxE * kVecThreads;
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
                for (int e = tid; e < npw * n; e += kVecThreads) sT[e] = src[e];
                cur_level = lf.level;
            }
            const double* fb = fvals + (size_t)lf.box * npbox * nd;
            for (int e = tid; e < npbox; e += kVecThreads) sf[e] = fb[(size_t)e * nd + d];
            for (int k = tid; k < npw; k += kVecThreads) spx[k] = cexpi(-ts[k] * lf.dx);
            for (int k = tid; k < H; k += kVecThreads) spy[k] = cexpi(-ts[k] * lf.dy);
            __syncthreads();
            for (int e = tid; e < n * H; e += kVecThreads) {
                int m1 = e % n, k2 = e / n;
                double sr = 0.0, si = 0.0;
                for (int m2 = 0; m2 < n; m2++) {
                    double f = sf[m1 + n * m2];
                    cplx t = sT[m2 * npw + k2];
                    sr = fma(t.x, f, sr);
                    si = fma(t.y, f, si);
                }
                sff[e] = cmul(cplx{sr, si}, spy[k2]);
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kUpMaxE; r++) {
                int e = e0 + tid + r * kVecThreads;
                if (e < nexp) {
                    int k1 = e % npw, k2 = e / npw;
                    const cplx* fcol = sff + n * k2;
                    cplx s{0.0, 0.0};
                    for (int m1 = 0; m1 < n; m1++) s = cfma(sT[m1 * npw + k1], fcol[m1], s);
                    acc[r] = cfma(s, spx[k1], acc[r]);
                }
            }
        }
        cplx* out = partial + ((size_t)ch.slot * nd + d) * nexp;
#pragma unroll
        for (int r = 0; r < kUpMaxE; r++) {
            int e = e0 + tid + r * kVecThreads;
            if (e < nexp) out[(size_t)(e % npw) * H + e / npw] = acc[r];
        }
    }
}

constexpr int kDownChunk = 8;  // y-frequency columns staged in LDS at a time

// Evaluation of the cutoff boxes' local expansions on the PW leaves, one
// workgroup per leaf.  With adj = 1 the same pass is the adjoint of the leaf
// transform: tpw is then the values-to-PW table tpoly ([level][m][k]), read
// conjugated and transposed as one table, and every y frequency has weight 1.
template <int NQ>
__global__ __launch_bounds__(kVecThreads) void k_down_vec(int n, int npw, int nd, int ntab, int adj,
                                                          const PwLeaf* __restrict__ leaves,
                                                          const cplx* __restrict__ tpw, const double* __restrict__ ts,
                                                          const cplx* __restrict__ loc, double* __restrict__ pot,
                                                          double* __restrict__ grad, double* __restrict__ hess,
                                                          double oscale, int accum) {
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
    if (adj) {
        const cplx* T = tpw + (size_t)lf.level * npw * n;  // [m][k]
        for (int e = tid; e < npw * n; e += kVecThreads) {
            const int k = e / n, i = e % n;
            const cplx v = T[(size_t)i * npw + k];
            sE[e] = cplx{v.x, -v.y};
        }
    } else {
        const cplx* E = tpw + (size_t)lf.level * 3 * npw * n;
        for (int e = tid; e < ntab * npw * n; e += kVecThreads) sE[e] = E[e];
    }
    for (int k = tid; k < npw; k += kVecThreads) spx[k] = cexpi(ts[k] * lf.dx);
    for (int k = tid; k < H; k += kVecThreads) spy[k] = cexpi(ts[k] * lf.dy);
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
            for (int e = tid; e < npw * mc; e += kVecThreads) {
                int m1 = e % npw, mm = e / npw;
                sL[mm * npw + m1] = cmul(L[(size_t)m1 * H + c0 + mm], spx[m1]);
            }
            __syncthreads();
            for (int e = tid; e < ntab * mc * n; e += kVecThreads) {
                int t = e / (mc * n), r = e % (mc * n);
                int mm = r / n, k1 = r % n;
                const cplx* Et = sE + t * npw * n;
                const cplx* Lc = sL + mm * npw;
                cplx s{0.0, 0.0};
                for (int m1 = 0; m1 < npw; m1++) s = cfma(Et[m1 * n + k1], Lc[m1], s);
                sff[(t * kDownChunk + mm) * n + k1] = cmul(s, spy[c0 + mm]);
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NQ; r++) {
                int q = tid + r * kVecThreads;
                if (q < npbox) {
                    int k1 = q % n, k2 = q / n;
                    for (int mm = 0; mm < mc; mm++) {
                        int m2 = c0 + mm;
                        double w = (adj || m2 < npw2) ? 1.0 : 0.5;
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
            int q = tid + r * kVecThreads;
            if (q < npbox) {
                size_t pb = (size_t)lf.box * npbox + q;
                const double w2 = 2.0 * oscale;
                auto put = [&](double& v, double x) { v = (accum ? v : 0.0) + w2 * x; };
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

// Adjoint of the evaluation on the vector kernels, one workgroup per (chunk,
// segment of the expansion) as k_up_vec: the cotangents g_o of every output
// channel o of the chunk's leaves, contracted with the conjugated evaluation
// tables of the channel (x table chan_x(o), y table chan_y(o)), the
// conjugated leaf phases and the weight 1/2 on the middle y frequency of an
// odd expansion, are summed into the chunk's partial expansion:
//   out[k1][k2] += w(k2) px(k1) py(k2) sum_o sum_{m1,m2} g_o[m1][m2]
//                  conj(E_x(o)[k1][m1]) conj(E_y(o)[k2][m2]).
template <int kUpMaxE>
__global__ __launch_bounds__(kVecThreads) void k_adj_eval_vec(int n, int npw, int nd, int ntab, LeafIn in,
                                                              const PwLeaf* __restrict__ leaves,
                                                              const UpChunk* __restrict__ chunks,
                                                              const cplx* __restrict__ tpw,
                                                              const double* __restrict__ ts,
                                                              cplx* __restrict__ partial) {
    extern __shared__ double smem[];
    const int H = (npw + 1) / 2, npw2 = npw / 2;
    const int nexp = npw * H;
    const int npbox = n * n;
    const int nout = ntab == 1 ? 1 : (ntab == 2 ? 3 : 6);
    cplx* sE = reinterpret_cast<cplx*>(smem);         // ntab*npw*n : [t][k][i]
    cplx* sff = sE + ntab * npw * n;                  // H*n : sff[k2*n + m1]
    cplx* spx = sff + n * H;                          // npw
    cplx* spy = spx + npw;                            // H (weights included)
    double* sg = reinterpret_cast<double*>(spy + H);  // n*n : one channel's cotangent
    const UpChunk ch = chunks[blockIdx.x];
    const int tid = threadIdx.x;
    const int e0 = blockIdx.y * kUpMaxE * kVecThreads;
    for (int d = 0; d < nd; d++) {
        cplx acc[kUpMaxE];
#pragma unroll
        for (int r = 0; r < kUpMaxE; r++) acc[r] = cplx{0.0, 0.0};
        int cur_level = -1;
        for (int li = 0; li < ch.count; li++) {
            const PwLeaf lf = leaves[ch.first + li];
            __syncthreads();
            if (lf.level != cur_level) {
                const cplx* E = tpw + (size_t)lf.level * 3 * npw * n;
                for (int e = tid; e < ntab * npw * n; e += kVecThreads) sE[e] = E[e];
                cur_level = lf.level;
            }
            for (int k = tid; k < npw; k += kVecThreads) spx[k] = cexpi(-ts[k] * lf.dx);
            for (int k = tid; k < H; k += kVecThreads) {
                const cplx v = cexpi(-ts[k] * lf.dy);
                const double w = (k < npw2) ? 1.0 : 0.5;
                spy[k] = cplx{w * v.x, w * v.y};
            }
            const size_t pb = (size_t)lf.box * npbox;
            for (int o = 0; o < nout; o++) {
                __syncthreads();  // sg and sff of the previous channel are consumed
                for (int e = tid; e < npbox; e += kVecThreads) sg[e] = leaf_in(in, o, pb + e, nd, d);
                __syncthreads();
                const cplx* Ex = sE + chan_x(o) * npw * n;
                const cplx* Ey = sE + chan_y(o) * npw * n;
                // y contraction
                for (int e = tid; e < n * H; e += kVecThreads) {
                    const int m1 = e % n, k2 = e / n;
                    double sr = 0.0, si = 0.0;
                    for (int m2 = 0; m2 < n; m2++) {
                        const double g = sg[m1 + n * m2];
                        const cplx t = Ey[k2 * n + m2];
                        sr = fma(t.x, g, sr);
                        si = fma(-t.y, g, si);
                    }
                    sff[e] = cmul(cplx{sr, si}, spy[k2]);
                }
                __syncthreads();
                // x contraction
#pragma unroll
                for (int r = 0; r < kUpMaxE; r++) {
                    const int e = e0 + tid + r * kVecThreads;
                    if (e < nexp) {
                        const int k1 = e % npw, k2 = e / npw;
                        const cplx* fcol = sff + n * k2;
                        const cplx* ex = Ex + k1 * n;
                        cplx s{0.0, 0.0};
                        for (int m1 = 0; m1 < n; m1++) s = cfma(cplx{ex[m1].x, -ex[m1].y}, fcol[m1], s);
                        acc[r] = cfma(s, spx[k1], acc[r]);
                    }
                }
            }
        }
        cplx* out = partial + ((size_t)ch.slot * nd + d) * nexp;
#pragma unroll
        for (int r = 0; r < kUpMaxE; r++) {
            const int e = e0 + tid + r * kVecThreads;
            if (e < nexp) out[(size_t)(e % npw) * H + e / npw] = acc[r];
        }
    }
}

// ---------------------------------------------------------------------------
// expansion-level stages
// ---------------------------------------------------------------------------

// Merge (bfgt/boxfgt.f:765-786 shifts the children's expansions to their
// parents level by level up to the cutoff; here the shifts are folded into
// the leaf phases and only the sum remains): every cutoff box with several
// partial expansions sums them into its mp slot.  boxes[i] = (box, first slot, count, 0), the partial sums
// in consecutive slots of the same buffer.  A workgroup serves kMergeCols
// consecutive coefficients of one box: its kMergeRows rows each add every
// kMergeRows-th partial sum (loads of different rows in flight together, so a
// box with many partial sums takes count / kMergeRows round trips), and the
// rows are added through LDS in a fixed order (deterministic results).
constexpr int kMergeCols = 32, kMergeRows = 8;
__global__ __launch_bounds__(kMergeCols * kMergeRows) void k_merge(int len, int ntile,
                                                                   const int4* __restrict__ boxes,
                                                                   cplx* __restrict__ mp) {
    __shared__ cplx part[kMergeRows][kMergeCols];
    const int4 g = boxes[blockIdx.x / ntile];
    const int col = threadIdx.x % kMergeCols, row = threadIdx.x / kMergeCols;
    const int e = (blockIdx.x % ntile) * kMergeCols + col;
    double r0 = 0, i0 = 0, r1 = 0, i1 = 0;
    if (e < len) {
        int c = row;
        // four partial sums in flight per thread
        for (; c + 3 * kMergeRows < g.z; c += 4 * kMergeRows) {
            const cplx a = mp[(size_t)(g.y + c) * len + e];
            const cplx b = mp[(size_t)(g.y + c + kMergeRows) * len + e];
            const cplx u = mp[(size_t)(g.y + c + 2 * kMergeRows) * len + e];
            const cplx v = mp[(size_t)(g.y + c + 3 * kMergeRows) * len + e];
            r0 += a.x + u.x;
            i0 += a.y + u.y;
            r1 += b.x + v.x;
            i1 += b.y + v.y;
        }
        for (; c < g.z; c += kMergeRows) {
            const cplx a = mp[(size_t)(g.y + c) * len + e];
            r0 += a.x;
            i0 += a.y;
        }
    }
    part[row][col] = cplx{r0 + r1, i0 + i1};
    __syncthreads();
    if (row == 0 && e < len) {
        cplx t = part[0][col];
#pragma unroll
        for (int k = 1; k < kMergeRows; k++) {
            t.x += part[k][col].x;
            t.y += part[k][col].y;
        }
        mp[(size_t)g.x * len + e] = t;
    }
}

// loc = mp(self) + sum over the PW list of shift(offset) * mp(source)
// (bfgt/boxfgt.f:805-822).
// The adjoint translation runs over the transposed lists with conjugated
// shifts (csign = -1): a translation by a pure phase is unitary.
__device__ inline cplx m2l_at(const cplx* __restrict__ wshift, const cplx* __restrict__ mp, int nexp, size_t len,
                              int a, int e, int j0, int j1, const int2* __restrict__ m2l, int q, double csign) {
    cplx s = mp[(size_t)a * len + e];
    for (int j = j0; j < j1; j++) {
        int2 sj = m2l[j];
        cplx w = wshift[(size_t)sj.y * nexp + q];
        w.y *= csign;
        s = cfma(w, mp[(size_t)sj.x * len + e], s);
    }
    return s;
}

__global__ void k_m2l(int nd, int nexp, const int* __restrict__ m2l_ptr, const int2* __restrict__ m2l,
                      const cplx* __restrict__ wshift, double csign, const cplx* __restrict__ mp,
                      cplx* __restrict__ loc) {
    const int a = blockIdx.x;
    const int len = nd * nexp;
    const int j0 = m2l_ptr[a], j1 = m2l_ptr[a + 1];
    for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < len; e += gridDim.y * blockDim.x)
        loc[(size_t)a * len + e] = m2l_at(wshift, mp, nexp, len, a, e, j0, j1, m2l, e % nexp, csign);
}

// Same translation, written mirror-folded (struct Fold) for the MFMA
// downward pass.  A workgroup serves G consecutive cutoff boxes (G = 4: four
// siblings, whose neighbours overlap) and reads each distinct source
// expansion once.  Only the H x H valid entries of the padded layout are
// written: the padding is zeroed once when the plan is built and never
// touched again.  Workgroups are mapped XCD-contiguously, so the expansions
// of neighbouring cutoff boxes are read from one L2.
template <int G>
__global__ void k_m2l_fold(int nd, int P, int gy, int nanc, const int* __restrict__ gptr,
                           const int2* __restrict__ glist, const int* __restrict__ patch,
                           const cplx* __restrict__ wshift, double csign, const cplx* __restrict__ mp,
                           Fold* __restrict__ fold) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int g = lb / gy;
    const int H = (P + 1) / 2, nexp = P * H;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const int len = nd * nexp;
    const int e = (lb % gy) * blockDim.x + threadIdx.x;
    if (e >= nd * H * H) return;
    const int d = e / (H * H), r = e % (H * H);
    const int m = r / H, k2 = r % H;
    const int mm = P - 1 - m;
    const int q1 = m * H + k2, q2 = mm * H + k2;
    const size_t o1 = (size_t)d * nexp + q1, o2 = (size_t)d * nexp + q2;
    // the shift to a neighbour offset (kx, ky) in {-1,0,1}^2 factors as
    // ex^kx(m) * ey^ky(k2) with ex = the (1,0) shift and ey = the (0,1) shift
    // (an offset of -1 is the conjugate): three loads replace one per partner
    const cplx ex1 = wshift[(size_t)5 * nexp + m * H], ex2 = wshift[(size_t)5 * nexp + mm * H];
    const cplx ey = wshift[(size_t)7 * nexp + k2];
    cplx l1[G], l2[G];
#pragma unroll
    for (int t = 0; t < G; t++) l1[t] = l2[t] = cplx{0.0, 0.0};
    // groups of four in patch form: the 4 x 4 cells of sources around the
    // 2 x 2 targets, column by column; per target the three sources of a
    // column in its list are summed with their y shifts, then the column sum
    // is shifted in x -- conj(ex)^dx conj(ey)^dy for a source at offset
    // (dx, dy) from its target, in the forward and the transposed lists alike
    const int* rec = G == 4 && patch != nullptr ? patch + (size_t)g * kPatchInts : nullptr;
    if (rec != nullptr && rec[0] != -2) {
        const cplx bx1 = {ex1.x, -ex1.y}, bx2 = {ex2.x, -ex2.y}, by = {ey.x, -ey.y};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            cplx v1[4], v2[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int sl = rec[4 * r + c];
                v1[r] = v2[r] = cplx{0.0, 0.0};
                if (sl >= 0) {
                    v1[r] = mp[(size_t)sl * len + o1];
                    v2[r] = mp[(size_t)sl * len + o2];
                }
            }
#pragma unroll
            for (int t = 0; t < G; t++) {
                const int tm = rec[16 + t];
                const int tx = (tm >> 9) & 1, ty = (tm >> 10) & 1;
                const int dx = c - tx - 1;
                if (dx < -1 || dx > 1) continue;
                cplx s1 = {0.0, 0.0}, s2 = {0.0, 0.0};
#pragma unroll
                for (int dy = -1; dy <= 1; dy++) {
                    if (!((tm >> ((dx + 1) + 3 * (dy + 1))) & 1)) continue;
                    // the source row ty + 1 + dy, with ty in {0, 1}
                    const cplx a1 = ty ? v1[dy + 2] : v1[dy + 1], a2 = ty ? v2[dy + 2] : v2[dy + 1];
                    if (dy == 0) {
                        s1 = cplx{s1.x + a1.x, s1.y + a1.y};
                        s2 = cplx{s2.x + a2.x, s2.y + a2.y};
                    } else {
                        const cplx w = dy > 0 ? by : cplx{by.x, -by.y};
                        s1 = cfma(w, a1, s1);
                        s2 = cfma(w, a2, s2);
                    }
                }
                if (dx == 0) {
                    l1[t] = cplx{l1[t].x + s1.x, l1[t].y + s1.y};
                    l2[t] = cplx{l2[t].x + s2.x, l2[t].y + s2.y};
                } else {
                    const cplx w1 = dx > 0 ? bx1 : cplx{bx1.x, -bx1.y}, w2 = dx > 0 ? bx2 : cplx{bx2.x, -bx2.y};
                    l1[t] = cfma(w1, s1, l1[t]);
                    l2[t] = cfma(w2, s2, l2[t]);
                }
            }
        }
    } else {
    // each distinct source of the group's G boxes is read once; the
    // expansions of four sources are loaded before any is used
    const int j0 = gptr[g], j1 = gptr[g + 1];
    auto add = [&](const int2 sj, const cplx v1, const cplx v2) {
#pragma unroll
        for (int t = 0; t < G; t++) {
            const int k = (sj.y >> (4 * t)) & 15;
            if (k == 15) continue;
            const int kx = k % 3 - 1, ky = k / 3 - 1;
            const cplx wy = ky == 0 ? cplx{1.0, 0.0} : cplx{ey.x, ky * ey.y};
            const cplx wx1 = kx == 0 ? cplx{1.0, 0.0} : cplx{ex1.x, kx * ex1.y};
            const cplx wx2 = kx == 0 ? cplx{1.0, 0.0} : cplx{ex2.x, kx * ex2.y};
            cplx w1 = cmul(wx1, wy), w2 = cmul(wx2, wy);
            w1.y *= csign;
            w2.y *= csign;
            l1[t] = cfma(w1, v1, l1[t]);
            l2[t] = cfma(w2, v2, l2[t]);
        }
    };
    int j = j0;
    for (; j + 4 <= j1; j += 4) {
        int2 sj[4];
        cplx v1[4], v2[4];
#pragma unroll
        for (int u = 0; u < 4; u++) sj[u] = glist[j + u];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            v1[u] = mp[(size_t)sj[u].x * len + o1];
            v2[u] = mp[(size_t)sj[u].x * len + o2];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) add(sj[u], v1[u], v2[u]);
    }
    for (; j < j1; j++) {
        const int2 sj = glist[j];
        add(sj, mp[(size_t)sj.x * len + o1], mp[(size_t)sj.x * len + o2]);
    }
    }
#pragma unroll
    for (int t = 0; t < G; t++) {
        const int a = G * g + t;
        if (a >= nanc) break;
        Fold f;
        if (mm != m) {
            f.S = cplx{l1[t].x + l2[t].x, l1[t].y + l2[t].y};
            f.D = cplx{l1[t].x - l2[t].x, l1[t].y - l2[t].y};
        } else {
            f.S = l1[t];
            f.D = cplx{0, 0};
        }
        fold[(size_t)a * nd * Mh4 * H16 + ((size_t)d * Mh4 + m) * H16 + k2] = f;
    }
}

// The local expansions of the cutoff boxes back in the plain layout
// [box][density][m1 * H + m2] from the mirror-folded one: L(m) = (S + D)/2,
// L(npw-1-m) = (S - D)/2, L = S on the middle frequency of an odd expansion.
__global__ void k_unfold(int nd, int P, int nanc, const Fold* __restrict__ fold, cplx* __restrict__ loc) {
    const int H = (P + 1) / 2, nexp = P * H;
    const int Mh4 = round_up(H, 4), H16 = round_up(H, 16);
    const long tot = (long)nanc * nd * H * H;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e % (H * H));
        const long ad = e / (H * H);
        const int m = r / H, k2 = r % H, mm = P - 1 - m;
        const Fold f = fold[ad * Mh4 * H16 + (long)m * H16 + k2];
        cplx* L = loc + ad * nexp;
        if (mm != m) {
            L[(long)m * H + k2] = cplx{0.5 * (f.S.x + f.D.x), 0.5 * (f.S.y + f.D.y)};
            L[(long)mm * H + k2] = cplx{0.5 * (f.S.x - f.D.x), 0.5 * (f.S.y - f.D.y)};
        } else {
            L[(long)m * H + k2] = f.S;
        }
    }
}

}  // namespace

// The cutoff boxes' local expansions of the last execution, [box][density]
// [nexp] (the vector kernels keep them so; the matrix-core path folds them).
int unfold_local(Plan& p, cplx* out) {
    if (!p.has_pw || p.nanc == 0) return BFGT_OK;
    if (!use_mfma(p)) {
        BFGT_HIP_CHECK(hipMemcpyAsync(out, p.d_loc, sizeof(cplx) * (size_t)p.nanc * p.nd * p.nexp,
                                      hipMemcpyDeviceToDevice, p.stream));
        return BFGT_OK;
    }
    const long tot = (long)p.nanc * p.nd * p.H * p.H;
    const int blocks = (int)std::min<long>((tot + 255) / 256, 4096);
    k_unfold<<<blocks, 256, 0, p.stream>>>(p.nd, p.npw, p.nanc, reinterpret_cast<const Fold*>(p.d_loc), out);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// pot/grad/hess += the cutoff boxes' local expansions (plain layout, as
// unfold_local leaves them) evaluated on an explicit list of boxes: box
// slot, level, cutoff box and scaled offset from it per entry.
int launch_pw_eval_list(Plan& p, const PwLeaf* list, int count, const cplx* loc, double* pot, double* grad,
                        double* hess) {
    if (count == 0 || !p.has_pw) return BFGT_OK;
    size_t sh = down_smem_bytes(p.n, p.npw, p.ifpgh);
    const int q = (p.n * p.n + kVecThreads - 1) / kVecThreads;
    auto run = [&](auto kern) -> int {
        if (sh > 65536)
            BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        kern<<<count, kVecThreads, sh, p.stream>>>(p.n, p.npw, p.nd, p.ifpgh, 0, list, p.d_tpw, p.d_ts, loc, pot,
                                                   grad, hess, 1.0, 1);
        BFGT_HIP_CHECK(hipGetLastError());
        return BFGT_OK;
    };
    if (q <= 1) return run(k_down_vec<1>);
    if (q <= 2) return run(k_down_vec<2>);
    return run(k_down_vec<4>);
}

bool use_mfma(const Plan& p) { return p.n <= 16 && (p.npw + 1) / 2 <= 16 * kUpMaxTM; }

size_t fold_len(const Plan& p) {
    int H = (p.npw + 1) / 2;
    return (size_t)p.nd * round_up(H, 4) * round_up(H, 16);
}

size_t loc_buffer_cplx(const Plan& p) {
    // the MFMA path stores the mirror-folded local expansion: 2 * Mh * H >= npw * H
    size_t H = (p.npw + 1) / 2;
    return std::max((size_t)p.nexp, 2 * H * H) * p.nd;
}

size_t up_smem_bytes(int n, int npw) {
    int H = (npw + 1) / 2;
    return sizeof(double) * n * n + sizeof(cplx) * ((size_t)npw * n + (size_t)n * H + npw + H);
}

size_t adj_eval_smem_bytes(int n, int npw, int ntab) {
    int H = (npw + 1) / 2;
    return sizeof(cplx) * ((size_t)ntab * npw * n + (size_t)n * H + npw + H) + sizeof(double) * n * n;
}

size_t down_smem_bytes(int n, int npw, int ntab) {
    int H = (npw + 1) / 2;
    return sizeof(cplx) * ((size_t)ntab * npw * n + (size_t)npw * kDownChunk + (size_t)ntab * kDownChunk * n + npw + H);
}

template <int E>
static int launch_up_vec(Plan& p, const double* fvals, size_t sh) {
    if (sh > 65536)
        BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)k_up_vec<E>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    const int nseg = (p.nexp + E * kVecThreads - 1) / (E * kVecThreads);
    k_up_vec<E><<<dim3(p.nchunks, nseg), kVecThreads, sh, p.stream>>>(p.n, p.npw, p.nd, fvals, p.d_pwleaf,
                                                                      p.d_chunks, p.d_tpoly, p.d_ts, p.d_mp);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// Upward MFMA launch, dispatched on NKS = ceil(n/4), NTM and NTAB.
struct UpArgs {
    LeafIn in;
    const cplx* tab;
    int lstride, ymid_half, ntab;
};

// one table, n <= 8: 2 x 2 leaf blocks over column pairs (k_up_quad);
// otherwise column
// sums (k_up_col), two leaves per pass with one table, one with more; with
// one table and at most two x tiles the chunks of a cutoff box are summed
// per workgroup (merge over workgroup pieces)
template <int NKS, int NTM, int NTAB>
static int launch_up_mfma3(Plan& p, const UpArgs& u, int nTN) {
    // chunks of one cutoff box summed per workgroup where there is a merge
    // to shrink, the accumulators fit LDS (two x tiles: 16 KB per wave) and
    // the registers leave room for it (one table; with the adjoint's two or
    // three tables the reducing form measured slower)
    const bool br = NTM <= 2 && NTAB == 1 && p.nmerge > 0;
    // densities on separate workgroups where the chunks alone give few waves
    // (one density: the same work without the density loop) and for the
    // adjoint's two or three tables (more work per wave); with many small
    // one-table chunks the per-wave costs would grow instead
    const bool ds = p.nd == 1 || NTAB > 1 || (long)nTN * p.nchunks < kUpSplitWaves;
    const int nblk = nTN * (ds ? p.nd : 1) * ((p.nchunks + kUpReduceWaves - 1) / kUpReduceWaves);
    if constexpr (NTAB == 1 && NKS <= 2) {
        // 2 x 2 leaf blocks (the forward transform, or the adjoint's one-table pass)
        auto kq = br ? (ds ? k_up_quad<NTM, NTM <= 2, true> : k_up_quad<NTM, NTM <= 2, false>)
                     : (ds ? k_up_quad<NTM, false, true> : k_up_quad<NTM, false, false>);
        kq<<<nblk, kWave * kUpReduceWaves, 0, p.stream>>>(p.n, p.npw, p.pstride, p.nd, nTN, p.nchunks, u.in.pot,
                                                            p.d_pwleaf, br ? p.d_upqccb : p.d_upqcc, p.d_upsegs,
                                                            p.d_upquads, u.tab, u.lstride, u.ymid_half, p.d_phase,
                                                            p.d_levbase, p.d_mp);
    } else {
        constexpr int LPC = 1;
        constexpr bool BRC = NTM <= 2 && NTAB == 1;
        auto kern = br ? (ds ? k_up_col<NKS, NTM, NTAB, LPC, BRC, true> : k_up_col<NKS, NTM, NTAB, LPC, BRC, false>)
                       : (ds ? k_up_col<NKS, NTM, NTAB, LPC, false, true> : k_up_col<NKS, NTM, NTAB, LPC, false, false>);
        kern<<<nblk, kWave * kUpReduceWaves, 0, p.stream>>>(p.n, p.npw, p.pstride, p.nd, nTN, p.nchunks, u.in,
                                                             p.d_pwleaf, br ? p.d_upccb : p.d_upcc, p.d_upcols, u.tab,
                                                             u.lstride, u.ymid_half, p.d_phase, p.d_levbase, p.d_mp);
    }
    p.merge_pieces = br;
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NKS, int NTM>
static int launch_up_mfma2(Plan& p, const UpArgs& u, int nTN) {
    switch (u.ntab) {
        case 1: return launch_up_mfma3<NKS, NTM, 1>(p, u, nTN);
        case 2: return launch_up_mfma3<NKS, NTM, 2>(p, u, nTN);
        default: return launch_up_mfma3<NKS, NTM, 3>(p, u, nTN);
    }
}

template <int NKS>
static int launch_up_mfma1(Plan& p, const UpArgs& u, int nTM, int nTN) {
    switch (nTM) {
        case 1: return launch_up_mfma2<NKS, 1>(p, u, nTN);
        case 2: return launch_up_mfma2<NKS, 2>(p, u, nTN);
        case 3: return launch_up_mfma2<NKS, 3>(p, u, nTN);
        default: return launch_up_mfma2<NKS, 4>(p, u, nTN);
    }
}

static int launch_up_mfma(Plan& p, const UpArgs& u) {
    int H = (p.npw + 1) / 2;
    int nTM = (H + 15) / 16, nTN = (H + 15) / 16;
    switch ((p.n + 3) / 4) {
        case 1: return launch_up_mfma1<1>(p, u, nTM, nTN);
        case 2: return launch_up_mfma1<2>(p, u, nTM, nTN);
        case 3: return launch_up_mfma1<3>(p, u, nTM, nTN);
        default: return launch_up_mfma1<4>(p, u, nTM, nTN);
    }
}

// densities on separate waves (nd) or in turn (1): separate where the
// chunks alone give few waves, as in the FP64 pass
static int split_density_waves(const Plan& p, int nTN) {
    return ((long)nTN * p.nchunks < kUpSplitWaves) ? p.nd : 1;
}

template <int NKS, int NTM>
static int launch_up_split2(Plan& p, const double* fvals, int nTN) {
    const int nds = split_density_waves(p, nTN);
    int nitems = p.nchunks * nds * nTN;
    int blocks = (nitems + kWavesPerBlock - 1) / kWavesPerBlock;
    k_up_split<NKS, NTM><<<blocks, kWave * kWavesPerBlock, 0, p.stream>>>(
        p.n, p.npw, p.pstride, p.nd, nds, nTN, nitems, fvals, p.d_pwleaf, p.d_upcc, p.d_upcols, p.d_tup_f,
        16 * p.pstride, p.d_phase, p.d_levbase, p.d_mp);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NKS>
static int launch_up_split1(Plan& p, const double* fvals, int nTM, int nTN) {
    switch (nTM) {
        case 1: return launch_up_split2<NKS, 1>(p, fvals, nTN);
        case 2: return launch_up_split2<NKS, 2>(p, fvals, nTN);
        case 3: return launch_up_split2<NKS, 3>(p, fvals, nTN);
        default: return launch_up_split2<NKS, 4>(p, fvals, nTN);
    }
}

template <int NTM>
static int launch_up_split_quad(Plan& p, const double* fvals, int nTN) {
    const int nds = split_density_waves(p, nTN);
    const int nitems = p.nchunks * nds * nTN;
    const int blocks = (nitems + kWavesPerBlock - 1) / kWavesPerBlock;
    k_up_split_quad<NTM><<<blocks, kWave * kWavesPerBlock, 0, p.stream>>>(
        p.n, p.npw, p.pstride, p.nd, nds, nTN, nitems, fvals, p.d_pwleaf, p.d_upqcc, p.d_upsegs, p.d_upquads, p.d_tup_f,
        16 * p.pstride, p.d_phase, p.d_levbase, p.d_mp);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}
static int launch_up_split(Plan& p, const double* fvals) {
    int H = (p.npw + 1) / 2;
    int nTM = (H + 15) / 16, nTN = (H + 15) / 16;
    if (p.n <= 8) {  // 2 x 2 leaf blocks, as the FP64 pass
        switch (nTM) {
            case 1: return launch_up_split_quad<1>(p, fvals, nTN);
            case 2: return launch_up_split_quad<2>(p, fvals, nTN);
            case 3: return launch_up_split_quad<3>(p, fvals, nTN);
            default: return launch_up_split_quad<4>(p, fvals, nTN);
        }
    }
    switch ((p.n + 3) / 4) {
        case 1: return launch_up_split1<1>(p, fvals, nTM, nTN);
        case 2: return launch_up_split1<2>(p, fvals, nTM, nTN);
        case 3: return launch_up_split1<3>(p, fvals, nTM, nTN);
        default: return launch_up_split1<4>(p, fvals, nTM, nTN);
    }
}

// chunks of one cutoff box summed per workgroup where there is a merge to
// shrink and the FP32 accumulators fit LDS (at most two x tiles), as in the
// FP64 pass
template <int NTM>
static int launch_up_b3_2(Plan& p, const double* fvals, int nTN) {
    const int nds = split_density_waves(p, nTN);
    const bool br = NTM <= 2 && p.nmerge > 0;
    const int blocks = nTN * nds * ((p.nchunks + kUpReduceWaves - 1) / kUpReduceWaves);
    auto kern = br ? k_up_b3<NTM, NTM <= 2> : k_up_b3<NTM, false>;
    kern<<<blocks, kWave * kUpReduceWaves, 0, p.stream>>>(p.n, p.npw, p.nd, nds, nTN, p.nchunks, fvals, p.d_pwleaf,
                                                          br ? p.d_upccb : p.d_upcc, p.d_upcols, p.nlevstart,
                                                          p.d_b3uy, p.d_b3ux, p.d_mp);
    p.merge_pieces = br;
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}
static int launch_up_b3(Plan& p, const double* fvals) {
    const int H = (p.npw + 1) / 2;
    const int nTM = (H + 15) / 16, nTN = (H + 15) / 16;
    switch (nTM) {
        case 1: return launch_up_b3_2<1>(p, fvals, nTN);
        case 2: return launch_up_b3_2<2>(p, fvals, nTN);
        case 3: return launch_up_b3_2<3>(p, fvals, nTN);
        default: return launch_up_b3_2<4>(p, fvals, nTN);
    }
}

int launch_upward(Plan& p, const double* fvals) {
    p.merge_pieces = false;
    if (p.nchunks == 0) return BFGT_OK;
    if (use_mfma(p) && p.precision == BFGT_PREC_SPLIT32) return launch_up_split(p, fvals);
    if (use_mfma(p) && p.precision == BFGT_PREC_BF16X3) return launch_up_b3(p, fvals);
    if (use_mfma(p)) {
        UpArgs u{LeafIn{fvals, nullptr, nullptr}, p.d_tup_f, 16 * p.pstride, 0, 1};
        return launch_up_mfma(p, u);
    }
    size_t sh = up_smem_bytes(p.n, p.npw);
    int e = (p.nexp + kVecThreads - 1) / kVecThreads;
    if (e <= 2) return launch_up_vec<2>(p, fvals, sh);
    if (e <= 4) return launch_up_vec<4>(p, fvals, sh);
    if (e <= 8) return launch_up_vec<8>(p, fvals, sh);
    return launch_up_vec<16>(p, fvals, sh);  // segments of 4096 entries
}

// the merge schedule of the upward pass that ran: per chunk, or per
// workgroup piece
int launch_reduce(Plan& p) {
    const int len = p.nd * p.nexp;
    const int ntile = (len + kMergeCols - 1) / kMergeCols;
    const int nbox = p.merge_pieces ? p.nmergeb : p.nmerge;
    if (nbox > 0) {
        k_merge<<<nbox * ntile, kMergeCols * kMergeRows, 0, p.stream>>>(
            len, ntile, p.merge_pieces ? p.d_mergeb : p.d_merge, p.d_mp);
        BFGT_HIP_CHECK(hipGetLastError());
    }
    return BFGT_OK;
}

static int launch_m2l_fold(Plan& p, const int* gptr, const int2* glist, const int* patch, double csign) {
    int H = (p.npw + 1) / 2;
    int gy = (p.nd * H * H + 255) / 256;
    const int G = p.m2l_group, ngroups = (p.nanc + G - 1) / G;
    auto kern = G == 4 ? k_m2l_fold<4> : k_m2l_fold<1>;
    kern<<<ngroups * gy, 256, 0, p.stream>>>(p.nd, p.npw, gy, p.nanc, gptr, glist, G == 4 ? patch : nullptr,
                                             p.d_wshift, csign, p.d_mp,
                                             reinterpret_cast<Fold*>(p.d_loc));
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_m2l(Plan& p) {
    if (use_mfma(p)) return launch_m2l_fold(p, p.d_m2lg_ptr, p.d_m2lg, p.d_m2lpatch, 1.0);
    int len = p.nd * p.nexp;
    int gy = (len + 255) / 256;
    k_m2l<<<dim3(p.nanc, gy), 256, 0, p.stream>>>(p.nd, p.nexp, p.d_m2l_ptr, p.d_m2l, p.d_wshift, 1.0, p.d_mp,
                                                  p.d_loc);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// adj = 1: the adjoint of the leaf transform (one table, tpoly), into pot
template <int NQ>
static int launch_down_vec(Plan& p, double* pot, double* grad, double* hess, size_t sh, int adj = 0) {
    if (sh > 65536)
        BFGT_HIP_CHECK(hipFuncSetAttribute((const void*)k_down_vec<NQ>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    k_down_vec<NQ><<<p.npwleaf, kVecThreads, sh, p.stream>>>(p.n, p.npw, p.nd, adj ? 1 : p.ifpgh, adj, p.d_pwleaf,
                                                              adj ? p.d_tpoly : p.d_tpw, p.d_ts, p.d_loc, pot, grad,
                                                              hess, p.out_scale, adj ? 1 : p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NTAB, int NTP>
static int launch_down_col(Plan& p, const cplx* tab, int lstride, int ymid_half, double* pot, double* grad,
                           double* hess) {
    int blocks = (p.ncolgrp + kWavesPerBlock - 1) / kWavesPerBlock;
    // leaves per y-contraction pass: four for one table; one for two or three
    // tables, which keeps the two-table kernel at three waves per SIMD (two
    // leaves per pass need two y-operand sets and two output batches: two
    // waves per SIMD, 10 % slower on the c2 tree)
    constexpr int LG = NTAB == 1 ? 4 : 1;
    k_down_col<NTAB, NTP, LG><<<blocks, kWave * kWavesPerBlock, 0, p.stream>>>(
        p.n, p.npw, p.pstride, p.nd, p.ncolgrp, p.d_colgrp, p.d_pwleaf, tab, lstride, ymid_half, p.d_phase,
        p.d_levbase, reinterpret_cast<const Fold*>(p.d_loc), pot, grad, hess, p.out_scale, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NTAB, int NTP>
static int launch_down_pair(Plan& p, const cplx* tab, int lstride, int ymid_half, double* pot, double* grad,
                            double* hess) {
    int blocks = (p.npair + kWavesPerBlock - 1) / kWavesPerBlock;
    k_down_pair<NTAB, NTP><<<blocks, kWave * kWavesPerBlock, 0, p.stream>>>(
        p.n, p.npw, p.pstride, p.nd, p.npair, p.d_pwpair, p.d_pwleaf, tab, lstride, ymid_half, p.d_phase,
        p.d_levbase, reinterpret_cast<const Fold*>(p.d_loc), pot, grad, hess, p.out_scale, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

template <int NTAB, int NTP>
static int launch_down_quad(Plan& p, const cplx* tab, int lstride, int ymid_half, double* pot, double* grad,
                            double* hess) {
    int blocks = (p.nqgroup * p.nd + kWavesPerBlock - 1) / kWavesPerBlock;
    k_down_quad<NTAB, NTP><<<blocks, kWave * kWavesPerBlock, 0, p.stream>>>(
        p.n, p.npw, p.pstride, p.nd, p.nqgroup, p.d_qgroups, p.d_quads, p.d_pwleaf, tab, lstride, ymid_half,
        p.d_phase, p.d_levbase, reinterpret_cast<const Fold*>(p.d_loc), pot, grad, hess, p.out_scale, p.accum);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// two y tiles per pass when their count is even; for n <= 8 2 x 2 leaf
// blocks over column pairs (leaf pairs where that leaves fewer than 1024
// waves), column groups above
template <int NTAB>
static int launch_down_mfma(Plan& p, const cplx* tab, int lstride, int ymid_half, double* pot, double* grad,
                            double* hess) {
    int nT = round_up((p.npw + 1) / 2, 16) / 16;
    if (p.n <= 8 && (long)p.nqgroup * p.nd < 1024) {
        if (nT % 2 == 0) return launch_down_pair<NTAB, 2>(p, tab, lstride, ymid_half, pot, grad, hess);
        return launch_down_pair<NTAB, 1>(p, tab, lstride, ymid_half, pot, grad, hess);
    }
    if (p.n <= 8) {
        if (nT % 2 == 0) return launch_down_quad<NTAB, 2>(p, tab, lstride, ymid_half, pot, grad, hess);
        return launch_down_quad<NTAB, 1>(p, tab, lstride, ymid_half, pot, grad, hess);
    }
    if (nT % 2 == 0) return launch_down_col<NTAB, 2>(p, tab, lstride, ymid_half, pot, grad, hess);
    return launch_down_col<NTAB, 1>(p, tab, lstride, ymid_half, pot, grad, hess);
}

template <int NTAB>
static int launch_down_split(Plan& p, double* pot, double* grad, double* hess) {
    const int nT = round_up((p.npw + 1) / 2, 16) / 16;
    if (p.n <= 8) {  // 2 x 2 leaf blocks over column pairs, one wave per (group, density)
        const int qblocks = (p.nqgroup * p.nd + kWavesPerBlock - 1) / kWavesPerBlock;
        auto runq = [&](auto kern) -> int {
            kern<<<qblocks, kWave * kWavesPerBlock, 0, p.stream>>>(
                p.n, p.npw, p.pstride, p.nd, p.nqgroup, p.d_qgroups, p.d_quads, p.d_pwleaf, p.d_tdn_f,
                3 * 16 * p.pstride, p.d_phase, p.d_levbase, reinterpret_cast<const Fold*>(p.d_loc), pot, grad, hess,
                p.out_scale, p.accum);
            BFGT_HIP_CHECK(hipGetLastError());
            return BFGT_OK;
        };
        switch (nT) {
            case 1: return runq(k_down_split_quad<NTAB, 1>);
            case 2: return runq(k_down_split_quad<NTAB, 2>);
            case 3: return runq(k_down_split_quad<NTAB, 3>);
            default: return runq(k_down_split_quad<NTAB, 4>);
        }
    }
    int blocks = (p.ncolgrp_short + kWavesPerBlock - 1) / kWavesPerBlock;
    auto run = [&](auto kern) -> int {
        kern<<<blocks, kWave * kWavesPerBlock, 0, p.stream>>>(
            p.n, p.npw, p.pstride, p.nd, p.ncolgrp_short, p.d_colgrp_short, p.d_pwleaf, p.d_tdn_f,
            3 * 16 * p.pstride, p.d_phase, p.d_levbase, reinterpret_cast<const Fold*>(p.d_loc), pot, grad, hess,
            p.out_scale, p.accum);
        BFGT_HIP_CHECK(hipGetLastError());
        return BFGT_OK;
    };
    switch (nT) {
        case 1: return run(k_down_split<NTAB, 1>);
        case 2: return run(k_down_split<NTAB, 2>);
        case 3: return run(k_down_split<NTAB, 3>);
        default: return run(k_down_split<NTAB, 4>);
    }
}

template <int NTAB>
static int launch_down_b3(Plan& p, double* pot, double* grad, double* hess) {
    const int nT = round_up((p.npw + 1) / 2, 16) / 16;
    const int blocks = (p.ncolgrp + kWavesPerBlock - 1) / kWavesPerBlock;
    auto run = [&](auto kern) -> int {
        kern<<<blocks, kWave * kWavesPerBlock, 0, p.stream>>>(
            p.n, p.npw, p.nd, p.ncolgrp, p.d_colgrp, p.d_pwleaf, p.nlevstart, p.b3nkb, p.d_b3dx, p.d_b3dy,
            reinterpret_cast<const Fold*>(p.d_loc), pot, grad, hess, p.out_scale, p.accum);
        BFGT_HIP_CHECK(hipGetLastError());
        return BFGT_OK;
    };
    switch (nT) {
        case 1: return run(k_down_b3<NTAB, 1>);
        case 2: return run(k_down_b3<NTAB, 2>);
        case 3: return run(k_down_b3<NTAB, 3>);
        default: return run(k_down_b3<NTAB, 4>);
    }
}

int launch_downward(Plan& p, double* pot, double* grad, double* hess) {
    if (p.npwleaf == 0) return BFGT_OK;
    if (use_mfma(p) && p.precision == BFGT_PREC_BF16X3) {
        if (p.ifpgh == 1) return launch_down_b3<1>(p, pot, grad, hess);
        if (p.ifpgh == 2) return launch_down_b3<2>(p, pot, grad, hess);
        return launch_down_b3<3>(p, pot, grad, hess);
    }
    if (use_mfma(p) && p.precision == BFGT_PREC_SPLIT32) {
        if (p.ifpgh == 1) return launch_down_split<1>(p, pot, grad, hess);
        if (p.ifpgh == 2) return launch_down_split<2>(p, pot, grad, hess);
        return launch_down_split<3>(p, pot, grad, hess);
    }
    if (use_mfma(p)) {
        const int ls = 3 * 16 * p.pstride;
        if (p.ifpgh == 1) return launch_down_mfma<1>(p, p.d_tdn_f, ls, 1, pot, grad, hess);
        if (p.ifpgh == 2) return launch_down_mfma<2>(p, p.d_tdn_f, ls, 1, pot, grad, hess);
        return launch_down_mfma<3>(p, p.d_tdn_f, ls, 1, pot, grad, hess);
    }
    size_t sh = down_smem_bytes(p.n, p.npw, p.ifpgh);
    int q = (p.n * p.n + kVecThreads - 1) / kVecThreads;
    if (q <= 1) return launch_down_vec<1>(p, pot, grad, hess, sh);
    if (q <= 2) return launch_down_vec<2>(p, pot, grad, hess, sh);
    return launch_down_vec<4>(p, pot, grad, hess, sh);
}

// ---------------------------------------------------------------------------
// adjoint (transposed operator) of the PW stages
// ---------------------------------------------------------------------------

// cotangents of pot/grad/hess at PW leaves -> per-chunk partial sums of the
// cutoff-level local-expansion cotangent
template <int E>
static int launch_adj_eval_vec(Plan& p, const LeafIn& in) {
    const size_t sh = adj_eval_smem_bytes(p.n, p.npw, p.ifpgh);
    if (sh > 65536)
        BFGT_HIP_CHECK(
            hipFuncSetAttribute((const void*)k_adj_eval_vec<E>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    const int nseg = (p.nexp + E * kVecThreads - 1) / (E * kVecThreads);
    k_adj_eval_vec<E><<<dim3(p.nchunks, nseg), kVecThreads, sh, p.stream>>>(
        p.n, p.npw, p.nd, p.ifpgh, in, p.d_pwleaf, p.d_chunks, p.d_tpw, p.d_ts, p.d_mp);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

int launch_adj_eval(Plan& p, const double* gpot, const double* ggrad, const double* ghess) {
    p.merge_pieces = false;
    if (p.nchunks == 0) return BFGT_OK;
    if (!use_mfma(p)) {
        const LeafIn in{gpot, ggrad, ghess};
        const int e = (p.nexp + kVecThreads - 1) / kVecThreads;
        if (e <= 2) return launch_adj_eval_vec<2>(p, in);
        if (e <= 4) return launch_adj_eval_vec<4>(p, in);
        if (e <= 8) return launch_adj_eval_vec<8>(p, in);
        return launch_adj_eval_vec<16>(p, in);  // segments of 4096 entries
    }
    UpArgs u{LeafIn{gpot, ggrad, ghess}, p.d_tup_a, 3 * 16 * p.pstride, 1, p.ifpgh};
    return launch_up_mfma(p, u);
}

// local-expansion cotangent (in d_mp) -> multipole cotangent in d_loc
// (mirror-folded for the matrix-core kernels)
int launch_adj_m2l(Plan& p) {
    if (use_mfma(p)) return launch_m2l_fold(p, p.d_m2lgT_ptr, p.d_m2lgT, p.d_m2lpatchT, -1.0);
    const int len = p.nd * p.nexp;
    const int gy = (len + 255) / 256;
    k_m2l<<<dim3(p.nanc, gy), 256, 0, p.stream>>>(p.nd, p.nexp, p.d_m2lT_ptr, p.d_m2lT, p.d_wshift, -1.0, p.d_mp,
                                                  p.d_loc);
    BFGT_HIP_CHECK(hipGetLastError());
    return BFGT_OK;
}

// The phased, sliced tables of the three-slice bf16 passes (k_b3_tables):
// every cell of every level from the cutoff level down.
int build_b3_tables(Plan& p) {
    if (!p.has_pw || !use_mfma(p)) return BFGT_OK;
    const int H = (p.npw + 1) / 2;
    const int nT = (H + 15) / 16;
    p.b3nkb = (round_up(H, 4) + 15) / 16;
    const int ncell = (1 << (p.nlevels - p.nlevstart + 1)) - 1;
    const int ntab = p.ifpgh;
    const size_t per = (size_t)6 * 64;  // v4s per entry
    const size_t n_dx = (size_t)ncell * ntab * p.b3nkb * per, n_dy = (size_t)ncell * ntab * nT * per;
    const size_t n_uy = (size_t)ncell * nT * per, n_ux = (size_t)ncell * nT * per;
    auto alloc = [&](v4s*& dst, size_t count) -> int {
        BFGT_HIP_CHECK(hipMalloc(&dst, sizeof(v4s) * count));
        p.allocations.push_back(dst);
        return BFGT_OK;
    };
    int rc;
    if ((rc = alloc(p.d_b3dx, n_dx)) || (rc = alloc(p.d_b3dy, n_dy)) || (rc = alloc(p.d_b3uy, n_uy)) ||
        (rc = alloc(p.d_b3ux, n_ux)))
        return rc;
    const int kinds[4][2] = {{ntab, p.b3nkb}, {ntab, nT}, {1, nT}, {1, nT}};
    v4s* outs[4] = {p.d_b3dx, p.d_b3dy, p.d_b3uy, p.d_b3ux};
    for (int k = 0; k < 4; k++) {
        const long tot = (long)ncell * kinds[k][0] * kinds[k][1] * 64;
        const int blocks = (int)std::min<long>((tot + 255) / 256, 4096);
        k_b3_tables<<<blocks, 256, 0, p.stream>>>(k, p.npw, p.pstride, p.nlevstart, ncell, kinds[k][0], kinds[k][1],
                                                   p.d_tup_f, p.d_tdn_f, p.d_phase, p.d_levbase, outs[k]);
        BFGT_HIP_CHECK(hipGetLastError());
    }
    return BFGT_OK;
}

// multipole cotangent -> leaf-value cotangent at PW leaves (accumulated)
int launch_adj_transform(Plan& p, double* fbar) {
    if (p.npwleaf == 0) return BFGT_OK;
    if (!use_mfma(p)) {
        const size_t sh = down_smem_bytes(p.n, p.npw, 1);
        const int q = (p.n * p.n + kVecThreads - 1) / kVecThreads;
        if (q <= 1) return launch_down_vec<1>(p, fbar, nullptr, nullptr, sh, 1);
        if (q <= 2) return launch_down_vec<2>(p, fbar, nullptr, nullptr, sh, 1);
        return launch_down_vec<4>(p, fbar, nullptr, nullptr, sh, 1);
    }
    return launch_down_mfma<1>(p, p.d_tdn_a, 16 * p.pstride, 0, fbar, nullptr, nullptr);
}


}  // namespace bfgt
