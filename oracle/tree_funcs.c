/*
 * Test functions for the tree builder and adapter.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference's tree routines (vol_tree_mem2, vol_tree_build2,
 * vol_tree_adap_fv, common/tree_ext.f:996, :1363, :2337) sample the user
 * function through an `external fun` with the calling sequence
 * fun(nd, xy, dpars, zpars, ipars, f), every argument by reference, one
 * point per call (common/tree_vol_coeffs_nd.f:537, :823).  oracle_tree_fun has
 * that sequence, so its address can be handed to the compiled reference;
 * oracle_tree_fun_eval evaluates the same function at many points, so that
 * the device tree code can be fed bit-identical values.
 *
 * The function: density d (0-based) of
 *     f_d(x, y) = (1 + d/2) * sum_k a_k exp(-((x - cx_k - vx_k t)^2 + (y - cy_k - vy_k t)^2) / s_k)
 * with dpars = [t, nbump, (a, cx, cy, vx, vy, s) x nbump].
 */
#include <math.h>

static void tree_fun_point(int nd, double x, double y, const double* dpars, double* f) {
    const double t = dpars[0];
    const int nb = (int)dpars[1];
    double s = 0.0;
    for (int k = 0; k < nb; k++) {
        const double* b = dpars + 2 + 6 * k;
        const double dx = x - b[1] - b[3] * t;
        const double dy = y - b[2] - b[4] * t;
        s += b[0] * exp(-(dx * dx + dy * dy) / b[5]);
    }
    for (int d = 0; d < nd; d++) f[d] = (1.0 + 0.5 * d) * s;
}

void oracle_tree_fun(const int* nd, const double* xy, const double* dpars, const void* zpars, const int* ipars,
                     double* f) {
    (void)zpars;
    (void)ipars;
    tree_fun_point(*nd, xy[0], xy[1], dpars, f);
}

/* xy (2, npts) point-major, f (nd, npts) point-major */
void oracle_tree_fun_eval(int nd, int npts, const double* xy, const double* dpars, double* f) {
    for (int i = 0; i < npts; i++) tree_fun_point(nd, xy[2 * i], xy[2 * i + 1], dpars, f + (long)nd * i);
}
