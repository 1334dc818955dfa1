/*
 * boxfgt_hip.h -- C ABI of the MI355X box fast Gauss transform.
 *
 * Plain pointers, sizes and status codes; no device or framework types in
 * the signatures.  Arrays use the reference's Fortran layouts (first index
 * fastest):
 *   itree(ltree) int32, iptr(8) int32 (1-based offsets into itree),
 *   centers(2,nboxes), boxsize(0:nlevels), fvals(nd,npbox,nboxes),
 *   pot(nd,npbox,nboxes), grad(nd,2,npbox,nboxes), hess(nd,3,npbox,nboxes)
 *   with hess ordered xx, xy, yy.  Outputs are accumulated (+=) on leaf boxes
 *   only, exactly like the reference.  All functions return BFGT_OK (0) or
 *   one of the BFGT_ERR_* codes; bfgt_last_error() describes the last failure.
 */
#ifndef BOXFGT_HIP_H
#define BOXFGT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BFGT_OK 0
#define BFGT_ERR_NDIM 1      /* ndim != 2 */
#define BFGT_ERR_IPOLY 2     /* ipoly not 0 (Legendre) or 1 (Chebyshev) */
#define BFGT_ERR_IFPGH 3     /* ifpgh not in 1..3 */
#define BFGT_ERR_IFNEWTREE 4 /* ifnewtree not 0 or 1, or 1 through an entry that cannot write the tree */
#define BFGT_ERR_TARGETS 5   /* ifpghtarg outside 0..3 */
#define BFGT_ERR_NORDER 6    /* norder < 1, norder > 32, or npbox != norder**2 */
#define BFGT_ERR_NPW 7       /* plane-wave length above 200 (the reference's ws/ts(200)) */
#define BFGT_ERR_TREE 8      /* inconsistent tree arrays */
#define BFGT_ERR_ARG 9       /* null pointer or invalid scalar argument */
#define BFGT_ERR_HIP 10      /* HIP runtime failure */
#define BFGT_ERR_NODEVICE 11 /* no GPU visible */
#define BFGT_ERR_RESOURCE 12 /* kernel resource limits (LDS) exceeded for this size, or a tree outgrew its capacity */
#define BFGT_ERR_CALLBACK 13 /* a user function callback returned non-zero */

/* arithmetic of the leaf transforms and near-field contractions */
#define BFGT_PREC_FP64 0    /* all-FP64 (the drop-in) */
#define BFGT_PREC_SPLIT32 1 /* split-FP32 hi/lo products, FP64 weights/phases/accumulation */
#define BFGT_PREC_BF16X3 2  /* leaf transforms as three-slice bf16 products (FP32 accumulation), near field split-FP32 */

/*
 * One complete box FGT call on host arrays, synchronous.
 * Replaces: subroutine boxfgt(nd,ndim,delta,eps,ipoly,iperiod,norder,npbox,
 *   nboxes,nlevels,ltree,itree,iptr,centers,boxsize,fvals,ifpgh,pot,grad,
 *   hess,ifnewtree,ntarg,targs,ifpghtarg,pote,grade,hesse,timeinfo)
 *   -- bfgt/boxfgt.f:5-8 (arguments :91-109).
 * Side effects as in the reference: the colleague lists in itree are
 * recomputed (bfgt/boxfgt.f:521-528); timeinfo(1..20) is zeroed and filled
 * (slot 1 set-up, 2 leaf->PW, 3 merge, 4 translation, 5 split (fused into 6),
 * 6 PW->pot/grad/hess, 7 near field, 12 extra targets; in slots the reference
 * leaves zero: 14 the host part of a new plan (lists, schedules), 15 its
 * table generation, 16 the rest of slot 1 (copies in, checks), 17 the copies
 * out, 18 = 1 if the call reused a kept plan).  ifnewtree must be
 * 0 here (BFGT_ERR_IFNEWTREE otherwise: the refining call writes nlevels and
 * centers, which this signature passes read-only; see bfgt_hip_boxfgt).  With ifpghtarg = 1..3 and ntarg > 0 the leaf pot (and grad, hess) are
 * interpolated at targs(2,ntarg) into pote(nd,ntarg), grade(nd,2,ntarg),
 * hesse(nd,3,ntarg), which are overwritten (bfgtmain STEP 11,
 * bfgt/boxfgt.f:1494-1563); channels above ifpgh are computed into scratch
 * arrays and never written to grad/hess (bfgt/boxfgt.f:207-277).
 * `precision` selects BFGT_PREC_FP64, BFGT_PREC_SPLIT32 or BFGT_PREC_BF16X3.
 * The reduced precisions apply to the matrix-core leaf transforms and near
 * field (norder <= 16, npw <= 128; BF16X3 runs the near field split-FP32); a
 * plan outside those sizes runs FP64 and reports precision 0, and a BF16X3
 * plan whose per-cell tables would exceed 256 MB runs split-FP32 and reports
 * 1 (bfgt_plan_query).
 * The plan built for a call (schedules, tables, device buffers) is kept and
 * reused by later calls with the same tree arrays (laddr, levels, parents,
 * children, centres, box sizes), parameters and device -- the drivers call
 * boxfgt on one tree several times per step -- so such a call costs the
 * copies and the kernels only.  The host arrays are copied through pinned
 * staging buffers (set up on the first call on a device) by several host
 * threads; a call that builds a plan copies its inputs on a second thread
 * while it builds the plan.  At most 4 plans are kept (least recently used
 * dropped); the environment variable BFGT_HOST_PLAN_CACHE=0 switches keeping
 * off.  Calls are serialised by a lock.
 */
int bfgt_hip_eval(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int npbox,
                  int nboxes, int nlevels, int ltree, int *itree, const int *iptr, const double *centers,
                  const double *boxsize, const double *fvals, int ifpgh, double *pot, double *grad, double *hess,
                  int ifnewtree, int ntarg, const double *targs, int ifpghtarg, double *pote, double *grade,
                  double *hesse, double *timeinfo, int precision);

/*
 * The complete boxfgt contract, nlevels and centers written back: the same
 * call as bfgt_hip_eval (arguments, kept plans, staging, status codes) plus
 * ifnewtree = 1.
 * Replaces: subroutine boxfgt(...) -- bfgt/boxfgt.f:5-8, every argument by
 * reference as there; with ifnewtree=1 the tree is adapted to pot after the
 * transform (bfgtmain STEPS 8-10, bfgt/boxfgt.f:1172-1483):
 *  - STEP 8: leaves above the finest level whose pot is not resolved
 *    (l2 tail of its tensor coefficients over eps * the l2 norm of pot) are
 *    split; the children's pot/grad/hess are evaluated from the cutoff-level
 *    local expansions and the parent's near-field list with tables for
 *    targets refined up to 3 times, and split in turn within the reference's
 *    generation limits;
 *  - STEP 9 (when the refinement added at least a tenth of the boxes):
 *    boxes whose four leaf children's pot, interpolated onto the parent, is
 *    resolved are merged, grad/hess interpolated too;
 *  - STEP 10: colleague lists, then 2:1 level restriction with the values of
 *    new boxes interpolated from their parents.
 * itree (all lists), centers, *nlevels (it may fall when coarsening empties
 * the finest level) and pot/grad/hess of every box of the new tree are
 * written, the boxes numbered as the reference numbers them; extra targets
 * are then located in the new tree.  nboxes is the capacity of the box
 * arrays (iptr's layout of itree must hold as many); a tree that would
 * outgrow it returns BFGT_ERR_RESOURCE with the caller's arrays partly
 * rewritten.  timeinfo slots 9-11 time the three steps, slot 19 counts the
 * boxes the refinement created and slot 20 the leaves at the finest level
 * that asked to be split: they are not split (the reference would create a
 * level that boxsize(0:nlevels) and laddr cannot hold).
 */
int bfgt_hip_boxfgt(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int npbox,
                    int nboxes, int *nlevels, int ltree, int *itree, const int *iptr, double *centers,
                    const double *boxsize, const double *fvals, int ifpgh, double *pot, double *grad, double *hess,
                    int ifnewtree, int ntarg, const double *targs, int ifpghtarg, double *pote, double *grade,
                    double *hesse, double *timeinfo, int precision);

/* Free the plans, device buffers and pinned staging buffers kept by
 * bfgt_hip_eval (no counterpart in the reference, whose boxfgt keeps nothing
 * between calls). */
int bfgt_hip_release_plans(void);

/* Resident plan: the tree, lists and FP64 tables stay in device memory for
 * repeated calls with the same tree, delta and eps (the solvers call 2-3x per
 * time step on one tree).  Splits the reference call into its set-up
 * (bfgtmain STEP 0, bfgt/boxfgt.f:485-698) and its evaluation (steps 1-6,
 * bfgt/boxfgt.f:720-1071). */
typedef struct bfgt_plan *bfgt_plan_t;

/* Set-up.  Host tree arrays (itree's colleague lists are recomputed, as in
 * the reference).  `stream` is the hipStream_t every later execution runs on;
 * NULL selects the device's default stream. */
int bfgt_plan_create(bfgt_plan_t *plan, int nd, int ndim, double delta, double eps, int ipoly, int iperiod,
                     int norder, int nboxes, int nlevels, int ltree, int *itree, const int *iptr,
                     const double *centers, const double *boxsize, int ifpgh, int precision, void *stream);

/* Evaluation on device arrays (fvals, and pot/grad/hess which are
 * accumulated), asynchronous on the plan's stream; allocates nothing.
 * grad/hess may be NULL when ifpgh does not ask for them.  timeinfo, when
 * non-NULL, is filled after a stream synchronisation. */
int bfgt_plan_execute(bfgt_plan_t plan, const double *d_fvals, double *d_pot, double *d_grad, double *d_hess,
                      double *timeinfo);

/* bfgt_plan_execute with the evaluation's workspace (the cutoff-level mp
 * expansions and chunk partial sums, and the local expansions) taken from a
 * static device array of the library instead of the plan's own buffers, and
 * no allocation of any kind.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The array holds 64 MiB per GPU,
 * kept until the process ends; a plan whose workspace is larger is refused
 * with BFGT_ERR_RESOURCE before any device work.  All calls on one GPU share
 * the array, so calls that overlap on one GPU (several host threads or
 * streams) race; each call clears what it needs zeroed.  Same arguments,
 * outputs and stream as bfgt_plan_execute, bit for bit the same results. */
int bfgt_plan_execute_noalloc(bfgt_plan_t plan, const double *d_fvals, double *d_pot, double *d_grad,
                              double *d_hess, double *timeinfo);

/* Extra targets on device arrays: interpolate the leaf arrays pot (and grad,
 * hess for ifpghtarg >= 2, 3) -- normally the outputs of bfgt_plan_execute --
 * at d_targs(2,ntarg) into d_pote(nd,ntarg), d_grade(nd,2,ntarg),
 * d_hesse(nd,3,ntarg) (overwritten).  Each target is located in its leaf by
 * descending the tree with ties to the plus side and evaluated with the
 * leaf's tensor-polynomial interpolant.  Replaces the extra-target step of
 * bfgtmain (bfgt/boxfgt.f:1494-1563: treedata_trans_nd +
 * treedata_evalt_nd, common/tree_data_routs_nd.f:16-90, 710-806).
 * Asynchronous on the plan's stream; allocates nothing. */
int bfgt_plan_eval_targets(bfgt_plan_t plan, int ntarg, const double *d_targs, int ifpghtarg, const double *d_pot,
                           const double *d_grad, const double *d_hess, double *d_pote, double *d_grade,
                           double *d_hesse);

/* Adjoint (transposed operator) of bfgt_plan_execute, on device arrays,
 * asynchronous on the plan's stream; allocates nothing.  No counterpart in
 * the reference: it exists for callers that differentiate through the
 * transform (gradient checks, adjoint solves).  With A the real-linear map
 * fvals -> (pot, grad, hess) of the plan's ifpgh on leaf boxes,
 *   fbar += A^T (gpot, ggrad, ghess)
 * in the plain Euclidean inner product of the arrays as laid out above
 * (entries of non-leaf boxes are neither read nor written).  ggrad/ghess may
 * be NULL when ifpgh does not ask for them.  Every stage is transposed:
 * conjugated, transposed PW tables; the translation over the transposed
 * lists with conjugated shifts; near-field tables read transposed per
 * (target, source) pair.  Arithmetic is FP64 for every plan precision.
 * Every plan has an adjoint, on the matrix-core kernels (norder <= 16, npw
 * <= 128) and on the vector kernels beyond them; BFGT_ERR_RESOURCE only where
 * the vector evaluation's adjoint would not fit the local data share (its
 * ifpgh evaluation tables and one leaf's cotangents of one channel at once).
 * timeinfo, when non-NULL, is filled like bfgt_plan_execute's, per
 * transposed stage. */
int bfgt_plan_adjoint(bfgt_plan_t plan, const double *d_gpot, const double *d_ggrad, const double *d_ghess,
                      double *d_fbar, double *timeinfo);

/* pot (+grad, hess) += scale * A fvals: bfgt_plan_execute with the outputs
 * scaled before they are accumulated.  The callers' time-quadrature loop
 * multiplies every boxfgt result by dt w_j / (4 pi (dt - t_j)) before adding
 * it to the volume potential (hpots/hpots2dadap.f:343-354); with one plan per
 * quadrature node this is one call per node into the same device array, with
 * no temporary and no host round trip.  Allocates nothing. */
int bfgt_plan_execute_scaled(bfgt_plan_t plan, double scale, const double *d_fvals, double *d_pot, double *d_grad,
                             double *d_hess, double *timeinfo);

/* pot (+grad, hess) = scale * A fvals on the leaf boxes: bfgt_plan_execute_scaled
 * with the leaf outputs overwritten instead of accumulated (entries of
 * non-leaf boxes are not touched).  No counterpart in the reference, whose
 * boxfgt adds into pot/grad/hess; its callers clear the output before every
 * call and add the scaled result into the volume potential afterwards
 * (hpots/hpots2dadap.f:320-354: vpotinc = 0, boxfgt, vpot += w vpotinc).
 * Here the evaluation writes every PW leaf's outputs once, without reading
 * them, and the leaves it does not reach (above the cutoff level, or all of
 * them in a plan without PW leaves) are cleared before the near field adds
 * into them.  The results equal bfgt_plan_execute_scaled's on zeroed
 * outputs.  Asynchronous on the plan's stream; allocates nothing. */
int bfgt_plan_execute_overwrite(bfgt_plan_t plan, double scale, const double *d_fvals, double *d_pot,
                                double *d_grad, double *d_hess, double *timeinfo);

/* Device coordinates d_xy(2, npbox, nboxes) of the tensor-grid nodes of every
 * leaf (zeros for non-leaf boxes), point i + n*j at centre + (x_i, x_j) *
 * boxsize/2 -- where the callers evaluate their forcing term
 * (hpots/hpots2dadap.f:296-306).  Asynchronous on the plan's stream. */
int bfgt_plan_leaf_points(bfgt_plan_t plan, double *d_xy);

int bfgt_plan_destroy(bfgt_plan_t plan);

/* Per-stage device timing over a window of executions.  After
 * bfgt_plan_profile_begin(plan, max_steps) the next max_steps executions record
 * HIP events around each stage on the plan's stream (the same stream the
 * kernels run on); bfgt_plan_profile_end synchronises and returns the summed
 * seconds per stage: [0] leaf values -> PW (merge folded in), [1] merge of
 * partial expansions, [2] mp -> loc translation, [3] split (0: folded into
 * [4]), [4] loc -> pot/grad/hess, [5] near field -- the timeinfo(2..7) slots of
 * the reference (bfgt/boxfgt.f:748-1076) -- and the number of executions seen.
 * bfgt_plan_profile_stage(plan, max_steps, stage) records only the two events
 * around one stage (0, 1, 2, 4 or 5; -1 = every stage, as profile_begin), so
 * that the other stages run back to back as in an unprofiled execution; the
 * other entries of stage_seconds are then 0.  Executions outside a profiling
 * window, and executions called without timeinfo, record no events at all. */
int bfgt_plan_profile_begin(bfgt_plan_t plan, int max_steps);
int bfgt_plan_profile_stage(bfgt_plan_t plan, int max_steps, int stage);
int bfgt_plan_profile_end(bfgt_plan_t plan, double *stage_seconds, int *nsteps);

/* Plan facts: ivals[0..15] = {npwlevel, npw, nexp, nlevstart, nlevend, nleafbox,
 * n_pw_leaves, n_chunks, n_ancestors, n_near_targets, n_near_pairs, has_pw,
 * has_near, iperiod_eff, nboxes, precision}; dvals[0..7] = {delta0, pmax,
 * t_plan_host_s, t_tables_s, matrix_core_path (1.0 / 0.0), n_translation_pairs,
 * n_near_sources, translation_group (cutoff boxes per translation workgroup:
 * 1 or 4)}. */
int bfgt_plan_query(bfgt_plan_t plan, int *ivals, double *dvals);

/* Work counts of the plan's schedule, per density (no counterpart in the
 * reference; they give the algorithmic FLOPs of the leaf passes):
 * counts[0] x contractions of the upward pass, [1] x contractions of the
 * downward pass, [2] PW leaves, [3] and [4] the same two counts for the
 * adjoint's evaluation and leaf transform, [5] expansions read and [6]
 * expansions written by the merge (a cutoff box whose leaves fit one upward
 * chunk is written by the upward pass directly).  A contraction serves a
 * whole column (PW leaves with one cutoff box, level and x cell) where the
 * kernels share it, else one leaf.  Entries past 7 are set to 0. */
int bfgt_plan_work(bfgt_plan_t plan, long long *counts, int ncounts);

/* Copy one device table of one level to host memory (parity tests of the FP64
 * table generation).  which: 0 values->PW (n x npw complex, [m][k]),
 * 1 PW->pot/d/dd (3 x npw x n complex, [t][k][i]), 2 near-field masked
 * tables (13 x 3 x n x n double, [k+6][t][j][m]), 3 near-field sparsity
 * bounds (13 x (n+1) x 2 int, 1-based), 4 M2L shifts (9 x npw x H complex,
 * [offset][x frequency][y frequency], H = (npw+1)/2), 5 leaf-to-ancestor
 * phases of one level (2^(level-nlevstart) x npw complex), 6 the cutoff-level
 * mp expansions of the last execution (n_ancestors x nd x npw x H complex;
 * `level` ignored), 7 its loc expansions as stored: on the matrix-core path
 * (dvals[4] = 1) mirror-folded pairs S = L(m,.) + L(npw-1-m,.),
 * D = L(m,.) - L(npw-1-m,.) (n_ancestors x nd x Mh4 x H16 x 2 complex, zero
 * padded, Mh4 = 4 ceil(H/4), H16 = 16 ceil(H/16)),
 * otherwise n_ancestors x nd x npw x H complex; 8 the near-field schedule's
 * CSR pointer over its target leaves (n_near_targets + 1 ints). */
int bfgt_plan_get_table(bfgt_plan_t plan, int which, int level, void *host_out, size_t nbytes);

/*
 * Adaptive quadtrees resident in device memory: the tree builder and the
 * per-step re-meshing of the solvers, on the GPU.
 *
 * A tree holds the reference's tree arrays (levels, parents, children,
 * colleagues, centres, box sizes) and one array of values on the leaf nodes,
 * vals(nd,npbox,nboxes), in device memory with room for max_boxes boxes.  Its
 * boxes are ordered by level, and within a level the children of a level's
 * boxes follow their parents' order (child order mm, pm, mp, pp); the
 * reference orders the boxes that its level-restriction and re-meshing steps
 * add differently, so box numbers may differ from its trees while the boxes,
 * their links and their values agree.
 *
 * The user function is a callback fun(ctx, nd, npts, d_xy, d_vals, stream)
 * that writes f at npts points: d_xy(2,npts) and d_vals(nd,npts) are device
 * arrays; d_xy is complete when the callback is called, and the callback
 * enqueues its work on `stream` (the tree's stream) or completes it before
 * returning, and returns 0 (non-zero stops the call with BFGT_ERR_CALLBACK).  It replaces the reference's `external fun`,
 * fun(nd,xy,dpars,zpars,ipars,f), which is called one point at a time; the
 * parameters travel in ctx.
 */
typedef struct bfgt_tree *bfgt_tree_t;
typedef int (*bfgt_tree_fun_t)(void *ctx, int nd, int npts, const double *d_xy, double *d_vals, void *stream);

/* Build a tree resolving fun on [-boxlen/2, boxlen/2]^2 and keep fun's values
 * on its leaf nodes as the tree's values.
 * Replaces: vol_tree_mem2 + vol_tree_build2 (common/tree_ext.f:996, :1363;
 * refinement rule vol_tree_find_box_refine, fun_err,
 * common/tree_vol_coeffs_nd.f:689-943; level restriction vol_tree_fix_lr,
 * :1240).  A box at level l is split when the tail of its tensor-product
 * coefficients (iptype 0/1/2: max, l1, l2 norm over the tail mask of
 * tens_prod_get_rmask, common/polytens.f:1579) times (boxsize(l)/2)**eta over
 * the mask norm exceeds eps * rsc * rintl(l) (rsc = 1, 1/boxlen**2,
 * 1/boxlen for iptype 0, 1, 2).  With rintl_in = NULL the norms rintl(l) are
 * computed while the tree grows, as vol_tree_mem2 does; otherwise
 * rintl_in(0..max_levels) are used, as vol_tree_build2 does.  rintl_out
 * (max_levels+1 entries, may be NULL) receives the norms used, the last
 * computed one repeated past the deepest level.  The tree is then made
 * level-restricted (2:1 across edges and corners, periodic wrap for
 * iperiod = 1), fun sampled on the boxes this adds, and the colleague lists
 * computed.  stream = NULL uses a stream of the tree's own. */
int bfgt_tree_build(bfgt_tree_t *out, int nd, int ipoly, int iperiod, double eps, double boxlen, int norder,
                    int iptype, double eta, bfgt_tree_fun_t fun, void *ctx, const double *rintl_in,
                    int max_levels, int max_boxes, double *rintl_out, void *stream);

/* Re-mesh the tree for a new fun and carry the tree's values along.
 * Replaces: vol_tree_adap_fv (common/tree_ext.f:2337) with the tree's ipoly,
 * iptype and iperiod: leaves whose fun is not resolved (rule as above with
 * eta = 1 and eps * rsc * rintl(l)) are split, down to max_levels, and the
 * values interpolated onto the children; boxes whose four children are
 * leaves are merged (from the finest level up) when fun and the values,
 * interpolated from the children onto the parent's nodes, both have an l2
 * tail at most eps * rsc (vol_tree_coarsen_f2, common/tree_ext.f:2062); then
 * the tree is level-restricted with the values interpolated onto the added
 * boxes (vol_tree_fix_lr_interp, common/tree_vol_coeffs_nd.f:2571).
 * rintl holds max_levels+1 norms.  norder must be even.  On an error other
 * than BFGT_ERR_ARG the tree is left in an unspecified state. */
int bfgt_tree_adapt(bfgt_tree_t tree, double eps, const double *rintl, bfgt_tree_fun_t fun, void *ctx,
                    int max_levels);

/* A device tree from a tree in the reference layout (itree/iptr as the
 * reference's builders lay it out, boxes in any order with the root first,
 * any capacity layout) and values vals(nd,npbox,nboxes) (NULL: zeros), with
 * room for max_boxes boxes; for re-meshing a tree built elsewhere.  The boxes
 * are renumbered (breadth first, children in parent order) and the colleague
 * lists recomputed; links, levels and cells are checked (BFGT_ERR_TREE). */
int bfgt_tree_import(bfgt_tree_t *out, int nd, int ipoly, int iperiod, int norder, int iptype, double eta,
                     int nboxes, int nlevels, const int *itree, const int *iptr, const double *centers,
                     const double *boxsize, const double *vals, int max_boxes, void *stream);

/* Tree facts: ivals[0..9] = {nboxes, nlevels, ltree (= 2*(nlevels+1) +
 * 17*nboxes), n_leaves, max_boxes, nd, norder, ipoly, iperiod, iptype};
 * *d_vals (may be NULL) receives the device address of vals(nd,npbox,nboxes),
 * valid until the next build or adapt call on the tree. */
int bfgt_tree_query(bfgt_tree_t tree, int *ivals, double **d_vals);

/* Copy the tree to host memory in the reference layout: itree(ltree) with
 * laddr, ilevel, iparent, nchild, ichild, ncoll, coll in that order, iptr(8),
 * centers(2,nboxes), boxsize(0:nlevels), and (when vals is not NULL)
 * vals(nd,npbox,nboxes). */
int bfgt_tree_export(bfgt_tree_t tree, int *itree, int *iptr, double *centers, double *boxsize, double *vals);

/* A resident FGT plan on the tree as it is now (bfgt_plan_create on its
 * arrays); execute it with the tree's values from bfgt_tree_query. */
int bfgt_tree_plan_create(bfgt_plan_t *out, bfgt_tree_t tree, double delta, double eps, int ifpgh, int precision,
                          void *stream);

int bfgt_tree_destroy(bfgt_tree_t tree);

const char *bfgt_last_error(void);
int bfgt_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* BOXFGT_HIP_H */
