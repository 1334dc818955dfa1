/*
 * boxfgt_oracle.c -- CPU FP64 restatement of the reference box fast Gauss
 * transform (2-D, ifnewtree=0, ifpghtarg=0).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (heat2d_per_box_amd, HIP kernels behind include/boxfgt_hip.h)
 * never links or calls it.
 *
 * It follows the reference stage by stage, level by level, in the same
 * loop and summation order, so that its output agrees with the Fortran
 * reference to rounding.  Each function cites the reference file:line it
 * restates.  All arithmetic is IEEE binary64.
 *
 * Conventions (mirroring the Fortran arrays, 0-based here):
 *   tree arrays live in itree at the 1-based offsets iptr[0..7]; box ids
 *   stored in itree are 1-based, -1 for "none";
 *   centers[2*b+i], fvals[(b*npbox+p)*nd+d], pot[(b*npbox+p)*nd+d],
 *   grad[((b*npbox+p)*2+i)*nd+d], hess[((b*npbox+p)*3+i)*nd+d];
 *   leaf grid point p = i + n*j has x index i (x fastest).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include <time.h>

typedef double complex cplx;

#define MAXN 64  /* largest polynomial order supported by the oracle */

static double wtime(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* 1-D polynomial families                                              */
/* ------------------------------------------------------------------ */

/* P_0..P_{n}(x) by the three-term recurrence (common/legeexps.f:597-634). */
void oracle_legepols(double x, int n, double *p) {
    p[0] = 1.0;
    if (n == 0) return;
    p[1] = x;
    for (int k = 1; k < n; k++)
        p[k + 1] = ((2 * k + 1) * x * p[k] - k * p[k - 1]) / (k + 1);
}

/* T_0..T_{n}(x) (common/chebexps.f:594-626). */
void oracle_chebpols(double x, int n, double *p) {
    p[0] = 1.0;
    if (n == 0) return;
    p[1] = x;
    for (int k = 1; k < n; k++) p[k + 1] = 2 * x * p[k] - p[k - 1];
}

/*
 * Gauss-Legendre nodes (ascending) and weights on [-1,1], and
 * u(i,j) = P_{i-1}(x_j) w_j (2i-1)/2, the values->coefficients matrix
 * (common/legeexps.f:140-220, nodes by legerts :2399).  Nodes are found by
 * Newton's method on P_n from Chebyshev-like first guesses; u is stored
 * column-major: u[i + n*j].
 */
void oracle_legeexps(int n, double *x, double *w, double *u) {
    double pols[MAXN + 2];
    int half = (n + 1) / 2;
    for (int i = 0; i < half; i++) {
        /* i-th root counted from the right end, then mirrored */
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 1; k < n; k++) {
                double p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16 && it > 2) break;
        }
        /* recompute derivative at the converged root */
        {
            double p0 = 1.0, p1 = z;
            for (int k = 1; k < n; k++) {
                double p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        double wt = 2.0 / ((1.0 - z * z) * dp * dp);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wt;
        w[i] = wt;
    }
    if (n % 2 == 1) x[half - 1] = 0.0;
    if (u) {
        for (int j = 0; j < n; j++) {
            oracle_legepols(x[j], n - 1, pols);
            for (int i = 0; i < n; i++) u[i + n * j] = pols[i] * w[j] * (2 * i + 1) / 2.0;
        }
    }
}

/*
 * Chebyshev (first kind) nodes cos((2i-1)pi/2n) in ascending order and the
 * discrete Chebyshev transform u(1,j) = 1/n, u(i,j) = 2/n T_{i-1}(x_j)
 * (common/chebexps.f:12-180).  Weights are the Fejer weights the reference
 * returns (sum over even powers).
 */
void oracle_chebexps(int n, double *x, double *w, double *u) {
    double h = M_PI / (2 * n);
    for (int i = 1; i <= n; i++) x[n - i] = cos((2 * i - 1) * h);
    double pols[MAXN + 2];
    for (int j = 0; j < n; j++) {
        oracle_chebpols(x[j], n - 1, pols);
        if (u)
            for (int i = 0; i < n; i++) u[i + n * j] = (i == 0 ? 1.0 : 2.0) * pols[i] / n;
        if (w) {
            double s = 2.0;
            for (int k = 2; k < n; k += 2) s -= 4.0 * pols[k] / ((double)(k + 1) * (k - 1));
            w[j] = s / n;
        }
    }
}

/* ------------------------------------------------------------------ */
/* plane-wave sizing and nodes (common/fgtterms.f)                      */
/* ------------------------------------------------------------------ */

/* fgtpwterms (common/fgtterms.f:72-141). */
void oracle_fgtpwterms(double bsize, double delta, double eps, int iperiod, double *pmax, int *npw) {
    if (iperiod == 0) {
        double rmax = bsize / sqrt(delta);
        double d = sqrt(log(1.0 / eps));
        *pmax = 2.0 * d;
        double h = 2 * M_PI / (rmax + d);
        *npw = 2 * ((int)(*pmax / h) + 1);
    } else {
        *npw = 2 * ((int)(bsize * sqrt(log(1.0 / eps) / delta) / M_PI) + 1) + 1;
    }
}

/* Midpoint rule t_j=(j+1/2)h, w_j=h/(2 sqrt(pi)) exp(-t_j^2/4), j=-npw/2..npw/2-1
 * (get_pwnodes itype=0, common/fgtterms.f:151-179). */
void oracle_pwnodes(double pmax, int npw, double *ws, double *ts) {
    int npw2 = npw / 2;
    double h = pmax / npw2;
    double w = h / (2.0 * sqrt(M_PI));
    for (int j = -npw2; j < npw2; j++) {
        ts[j + npw2] = (j + 0.5) * h;
        ws[j + npw2] = w * exp(-ts[j + npw2] * ts[j + npw2] / 4);
    }
}

/* Fourier-series nodes of the periodic Gaussian, t_j=2 pi j/L,
 * w_j = sqrt(pi delta)/L exp(-(h j sqrt(delta)/2)^2)
 * (get_periodic_pwnodes, common/fgtterms.f:189-209). */
void oracle_periodic_pwnodes(double bsize0, double delta, int npw, double *ws, double *ts) {
    int npw2 = npw / 2;
    double h = 2 * M_PI / bsize0;
    double w = sqrt(M_PI * delta) / bsize0;
    double hw = h * sqrt(delta) / 2;
    for (int j = -npw2; j <= (npw - 1) / 2; j++) {
        ts[j + npw2] = j * h;
        ws[j + npw2] = w * exp(-(hw * j) * (hw * j));
    }
}

/* ------------------------------------------------------------------ */
/* spherical Bessel functions j_0..j_{n-1}(x), x real (besseljs3d.f)  */
/* ------------------------------------------------------------------ */

/* Miller's backward recurrence, normalised by j_0 or j_1, whichever is
 * larger in magnitude (common/besseljs3d.f:5-168 restated for real x). */
void oracle_sph_besselj(int n, double x, double *fj) {
    if (fabs(x) < 1e-200) {
        fj[0] = 1.0;
        for (int i = 1; i < n; i++) fj[i] = 0.0;
        return;
    }
    double xinv = 1.0 / x;
    /* upward sweep from order n-1 until the recurrence blows up */
    double fm1 = 0.0, f0 = 1.0;
    int ntop = n + 100000;
    for (int i = n - 1 < 0 ? 0 : n - 1; i < n + 100000; i++) {
        double f1 = (2 * i + 1.0) * xinv * f0 - fm1;
        if (f1 * f1 > 1e40) { ntop = i + 1; break; }
        fm1 = f0;
        f0 = f1;
    }
    double *tmp = (double *)malloc(sizeof(double) * (ntop + 2));
    tmp[ntop] = 0.0;
    tmp[ntop - 1] = 1.0;
    for (int i = ntop - 1; i >= 1; i--) {
        tmp[i - 1] = (2 * i + 1.0) * xinv * tmp[i] - tmp[i + 1];
        if (tmp[i - 1] * tmp[i - 1] > 1e40) {
            /* rescale everything above to keep the tail in range */
            for (int k = i - 1; k <= ntop; k++) tmp[k] *= 1e-40 * 1e20;
        }
    }
    double j0 = sin(x) * xinv;
    double j1 = j0 * xinv - cos(x) * xinv;
    double sc = (fabs(j1) > fabs(j0)) ? j1 / tmp[1] : j0 / tmp[0];
    for (int i = 0; i < n; i++) fj[i] = tmp[i] * sc;
    free(tmp);
}

/* ------------------------------------------------------------------ */
/* 1-D tables (bfgt/bfgt_volrouts.f)                                    */
/* ------------------------------------------------------------------ */

/* values -> PW table tab(m,j) (column-major, n x npw) for one level,
 * mk_leg2pw_tables / mk_cheb2pw_tables (bfgt/bfgt_volrouts.f:58-275). */
void oracle_poly2pw(int n, int ipoly, int npw, const double *ws, const double *ts, double delta,
                    double boxdim, cplx *tab) {
    cplx *tmp = (cplx *)calloc((size_t)n * npw, sizeof(cplx));
    double fj[MAXN + 2];
    double fac = boxdim / (2 * sqrt(delta));
    double xl[MAXN], wl[MAXN], ul[MAXN * MAXN];
    double vcl[MAXN * MAXN]; /* Chebyshev->Legendre coefficient map v(k,i) */
    if (ipoly == 1) {
        oracle_legeexps(n, xl, wl, ul);
        double cp[MAXN + 2];
        double chebv[MAXN * MAXN];
        for (int i = 0; i < n; i++) {
            oracle_chebpols(xl[i], n - 1, cp);
            for (int k = 0; k < n; k++) chebv[k + n * i] = cp[k];
        }
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                double dd = 0;
                for (int k = 0; k < n; k++) dd += ul[j + n * k] * chebv[i + n * k];
                vcl[j + n * i] = dd;
            }
    }
    for (int j = npw / 2; j < npw; j++) {
        double dd = ws[j] * boxdim;
        if (fabs(ts[j]) < 1e-12) {
            if (ipoly == 0) {
                tmp[0 + n * j] = dd;
                for (int m = 1; m < n; m++) tmp[m + n * j] = 0;
            } else {
                for (int m = 0; m < n; m += 2) tmp[m + n * j] = dd / (1.0 - (double)m * m);
                for (int m = 1; m < n; m += 2) tmp[m + n * j] = 0;
            }
        } else {
            oracle_sph_besselj(n, ts[j] * fac, fj);
            cplx legv[MAXN];
            cplx eyem = 1.0;
            for (int m = 0; m < n; m++) {
                legv[m] = dd * fj[m] / eyem;
                eyem *= I;
            }
            if (ipoly == 0) {
                for (int m = 0; m < n; m++) tmp[m + n * j] = legv[m];
            } else {
                for (int i = 0; i < n; i++) {
                    cplx cd = 0;
                    for (int k = 0; k < n; k++) cd += vcl[k + n * i] * legv[k];
                    tmp[i + n * j] = cd;
                }
            }
        }
        int jm = npw - 1 - j;
        if (jm != j)
            for (int m = 0; m < n; m++) tmp[m + n * jm] = conj(tmp[m + n * j]);
    }
    double xq[MAXN], wq[MAXN], u[MAXN * MAXN];
    if (ipoly == 0) oracle_legeexps(n, xq, wq, u);
    else oracle_chebexps(n, xq, wq, u);
    for (int i = 0; i < npw; i++)
        for (int j = 0; j < n; j++) {
            cplx cd = 0;
            for (int k = 0; k < n; k++) cd += tmp[k + n * i] * u[k + n * j];
            tab[j + n * i] = cd;
        }
    free(tmp);
}

/* PW -> pot/grad/hess tables E(j,i)=exp(i t_j x_i D/(2 sqrt delta)), times
 * (i t_j/sqrt delta) per derivative (mk_pw2pgh_tables,
 * bfgt/bfgt_volrouts.f:339-401).  Column-major npw x n. */
void oracle_pw2pgh(int n, int npw, const double *ts, const double *xs, double delta, double boxdim,
                   cplx *tp, cplx *tx, cplx *txx) {
    double dsq = boxdim / 2 / sqrt(delta);
    cplx dfac = I * dsq * 2 / boxdim;
    int npw2 = npw / 2;
    for (int i = 0; i < n; i++) {
        double x = xs[i] * dsq;
        for (int j = npw2; j < npw; j++) {
            cplx e = cexp(I * ts[j] * x);
            cplx ex = ts[j] * e * dfac;
            cplx exx = ts[j] * ex * dfac;
            tp[j + npw * i] = e;
            tx[j + npw * i] = ex;
            txx[j + npw * i] = exx;
            int jm = npw - 1 - j;
            tp[jm + npw * i] = conj(e);
            tx[jm + npw * i] = conj(ex);
            txx[jm + npw * i] = conj(exx);
        }
    }
}

/* I_n(t) = int_{-1}^{1} P_n(x) exp(-(t-x)^2/lambda^2) dx by the five-term
 * recurrence, with first and second t-derivatives; the lambda and box
 * scalings match mk_legetab_recur (bfgt/bfgt_volrouts.f:1095-1187). */
static void legetab_recur(int m, double targ, double lambda, double boxdim, double *fint, double *fx,
                          double *fxx) {
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[MAXN + 2], fpp[MAXN + 2];
    double l2 = lambda * lambda;
    double a = (-1 - targ) / lambda, b = (1 - targ) / lambda;
    double ea = exp(-a * a), eb = exp(-b * b);
    double d1 = lambda * (eb - ea), d2 = lambda * (eb + ea);
    fint[0] = sqrtpih * (erf(b) - erf(a));
    if (m > 1) fint[1] = -d1 / 2 + targ * fint[0];
    if (m > 2) fint[2] = 0.75 * (-d2 - targ * d1 + (l2 - 2.0 / 3 + 2 * targ * targ) * fint[0]);
    if (m > 3)
        fint[3] = 5.0 * (-d1 + 2 * targ * (4.0 * fint[2] / 3 - fint[0] / 3) - (0.4 - 4 * l2) * fint[1]) / 8;
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

/* Chebyshev counterpart, mk_chebtab_recur (bfgt/bfgt_volrouts.f:1191-1294). */
static void chebtab_recur(int m, double targ, double lambda, double boxdim, double *fint, double *fx,
                          double *fxx) {
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[MAXN + 2], fpp[MAXN + 2];
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

/* Table workhorse mk_loctab (bfgt/bfgt_volrouts.f:920-1060): coefficient
 * tables tab(m,j) = int_src p_m e^{-(t_j - y)^2/delta} dy and t-derivatives,
 * source box centred at 0, targets tnodes/scale + tboxcen in source units. */
static void mk_loctab(int ipoly, int n, double delta, const double *tnodes, double tboxdim,
                      double tboxcen, double scale, double *tab, double *tabx, double *tabxx, int nquad,
                      const double *xnodes, const double *wts, const double *polyv) {
    double targs[MAXN];
    for (int i = 0; i < n; i++) targs[i] = tnodes[i] / scale + tboxcen;
    double bs = tboxdim * scale;
    double sigma = 4 * delta / (bs * bs);
    double lambda = sqrt(sigma);
    double dmax = sqrt(log(1 / 1e-16) * sigma);
    if (lambda <= 0.125) {
        double fint[MAXN + 2], fx[MAXN + 2], fxx[MAXN + 2];
        for (int j = 0; j < n; j++) {
            double t = targs[j], dis = 0;
            if (t < -1) dis = -1 - t;
            if (t > 1) dis = t - 1;
            if (dis > dmax) {
                for (int m = 0; m < n; m++) tab[m + n * j] = tabx[m + n * j] = tabxx[m + n * j] = 0;
            } else {
                if (ipoly == 0) legetab_recur(n, t, lambda, bs, fint, fx, fxx);
                else chebtab_recur(n, t, lambda, bs, fint, fx, fxx);
                for (int m = 0; m < n; m++) {
                    tab[m + n * j] = fint[m];
                    tabx[m + n * j] = fx[m];
                    tabxx[m + n * j] = fxx[m];
                }
            }
        }
    } else {
        double *wexp = (double *)malloc(sizeof(double) * nquad * n);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < nquad; i++) {
                double dx = targs[j] - xnodes[i];
                wexp[i + nquad * j] = (fabs(dx) < dmax) ? exp(-dx * dx / sigma) * wts[i] : 0.0;
            }
        double bsh = bs / 2;
        for (int j = 0; j < n; j++) {
            double t = targs[j], dis = 0;
            if (t < -1) dis = -1 - t;
            if (t > 1) dis = t - 1;
            if (dis > dmax) {
                for (int m = 0; m < n; m++) tab[m + n * j] = tabx[m + n * j] = tabxx[m + n * j] = 0;
                continue;
            }
            for (int m = 0; m < n; m++) {
                double r = 0, rx = 0, rxx = 0;
                for (int i = 0; i < nquad; i++) {
                    double dx = t - xnodes[i];
                    double dd = -2 * dx / sigma * 2 / bs;
                    double pw = polyv[i + nquad * m] * wexp[i + nquad * j];
                    r += pw;
                    rx += pw * dd;
                    rxx += pw * (dd * dd - 2 / delta);
                }
                tab[m + n * j] = r * bsh;
                tabx[m + n * j] = rx * bsh;
                tabxx[m + n * j] = rxx * bsh;
            }
        }
        free(wexp);
    }
}

/* compute_sparse_pattern (bfgt/bfgt_volrouts.f:1300-1378): per column the
 * first/last row above eps*max/n (1-based, 0/-1 when empty), and in slot n
 * the first/last non-empty column. */
static void sparse_pattern(double eps, int n, const double *a, int *ind) {
    for (int i = 0; i <= n; i++) { ind[2 * i] = 0; ind[2 * i + 1] = -1; }
    double dmax = 0;
    for (int i = 0; i < n * n; i++) if (fabs(a[i]) > dmax) dmax = fabs(a[i]);
    if (dmax <= 2e-16) return;
    double thr = eps * dmax / n;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) if (fabs(a[j + n * i]) > thr) { ind[2 * i] = j + 1; break; }
        for (int j = n - 1; j >= 0; j--) if (fabs(a[j + n * i]) > thr) { ind[2 * i + 1] = j + 1; break; }
    }
    for (int i = 0; i < n; i++) if (ind[2 * i] > 0) { ind[2 * n] = i + 1; break; }
    for (int i = n - 1; i >= 0; i--) if (ind[2 * i + 1] > 0) { ind[2 * n + 1] = i + 1; break; }
}

/* merge the patterns of tab/tabx/tabxx the way mk_loctab_all does
 * (bfgt/bfgt_volrouts.f:603-625). */
static void merge_patterns(int n, const int *ia, const int *ib, const int *ic, int *out) {
    for (int j = 0; j <= n; j++) {
        int i2 = ia[2 * j + 1];
        if (i2 < ib[2 * j + 1]) i2 = ib[2 * j + 1];
        if (i2 < ic[2 * j + 1]) i2 = ic[2 * j + 1];
        int i1 = ia[2 * j];
        if (ib[2 * j] != 0 && i1 > ib[2 * j]) i1 = ib[2 * j];
        if (ic[2 * j] != 0 && i1 > ic[2 * j]) i1 = ic[2 * j];
        if (i2 > 0 && i1 == 0) i1 = 1;
        out[2 * j] = i1;
        out[2 * j + 1] = i2;
    }
}

/* half-width nloctab = 2^(mrefinelev+1) (mrefinelev+3) of the offset range
 * of the near-field tables (bfgt/boxfgt.f:625): 6 for mrefinelev=0, 96 for
 * the refining call's mrefinelev=3. */
static int loctab_half(int mref) { return (1 << (mref + 1)) * (mref + 3); }

/*
 * All near-field tables of one target level, mk_loctab_all
 * (bfgt/bfgt_volrouts.f:407-660) for targets refined up to mref times.
 * Offsets k=-nloctab..nloctab (index k+nloctab, nloctab = loctab_half);
 * tables column-major n x n per offset: tab[(k+nloctab)*n*n + m + n*j]
 * (m = source node, j = target node); ind[(k+nloctab)*2*(n+1) + 2*j + {0,1}].
 * Offsets the enumeration never reaches stay zero with empty patterns.
 */
void oracle_loctab_level_m(double eps, int ipoly, int n, double delta, double boxdim, int mref, double *tab,
                           double *tabx, double *tabxx, int *ind) {
    const int nloctab = loctab_half(mref), m0 = mref + 1, nquad = 50;
    const int nslot = 2 * nloctab + 1;
    double tnodes[MAXN], ws[MAXN], umat[MAXN * MAXN];
    if (ipoly == 0) oracle_legeexps(n, tnodes, ws, umat);
    else oracle_chebexps(n, tnodes, ws, umat);
    double xn[50], wn[50];
    oracle_legeexps(nquad, xn, wn, NULL);
    double *polyv = (double *)malloc(sizeof(double) * nquad * n);
    double pt[MAXN + 2];
    for (int i = 0; i < nquad; i++) {
        if (ipoly == 0) oracle_legepols(xn[i], n - 1, pt);
        else oracle_chebpols(xn[i], n - 1, pt);
        for (int j = 0; j < n; j++) polyv[i + nquad * j] = pt[j];
    }
    size_t nn = (size_t)n * n;
    memset(tab, 0, sizeof(double) * nn * nslot);
    memset(tabx, 0, sizeof(double) * nn * nslot);
    memset(tabxx, 0, sizeof(double) * nn * nslot);
    for (int i = 0; i < nslot * 2 * (n + 1); i += 2) { ind[i] = 0; ind[i + 1] = -1; }
    double t0[MAXN * MAXN], t1[MAXN * MAXN], t2[MAXN * MAXN];
    int ia[2 * (MAXN + 1)], ib[2 * (MAXN + 1)], ic[2 * (MAXN + 1)];
    for (int i0 = m0; i0 >= -1; i0--) {
        for (int j0 = 1; j0 <= (1 << m0); j0++) {
            int k;
            if (i0 > 0) k = -(1 << (m0 - i0)) - (1 << (m0 - i0 + 1)) * (j0 - 1);
            else if (i0 == 0) k = -(1 << (m0 - i0 + 1)) * (j0 - 1);
            else k = -(1 << m0) - (1 << (m0 + 1)) * (j0 - 1);
            if (k < -nloctab) break;
            double scale = pow(2.0, i0);
            double cen = k / (double)(1 << m0);
            mk_loctab(ipoly, n, delta, tnodes, boxdim, cen, scale, t0, t1, t2, nquad, xn, wn, polyv);
            double *A = tab + (k + nloctab) * nn, *Ax = tabx + (k + nloctab) * nn, *Axx = tabxx + (k + nloctab) * nn;
            /* fold U: A = U^T * t (dgemm 't','n') */
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    double s0 = 0, s1 = 0, s2 = 0;
                    for (int l = 0; l < n; l++) {
                        double ul = umat[l + n * i];
                        s0 += ul * t0[l + n * j];
                        s1 += ul * t1[l + n * j];
                        s2 += ul * t2[l + n * j];
                    }
                    A[i + n * j] = s0;
                    Ax[i + n * j] = s1;
                    Axx[i + n * j] = s2;
                }
            /* mirror to the targets on the other side; done in place, in the
             * reference's loop order, so k=0 is symmetrised the same way */
            double *B = tab + (-k + nloctab) * nn, *Bx = tabx + (-k + nloctab) * nn,
                   *Bxx = tabxx + (-k + nloctab) * nn;
            for (int j = 0; j < n; j++)
                for (int m = 0; m < n; m++) {
                    B[m + n * j] = A[(n - 1 - m) + n * (n - 1 - j)];
                    Bx[m + n * j] = -Ax[(n - 1 - m) + n * (n - 1 - j)];
                    Bxx[m + n * j] = Axx[(n - 1 - m) + n * (n - 1 - j)];
                }
            int kk[2] = {-k, k};
            for (int s = 0; s < 2; s++) {
                int q = kk[s] + nloctab;
                sparse_pattern(eps, n, tab + q * nn, ia);
                sparse_pattern(eps, n, tabx + q * nn, ib);
                sparse_pattern(eps, n, tabxx + q * nn, ic);
                merge_patterns(n, ia, ib, ic, ind + q * 2 * (n + 1));
            }
        }
    }
    free(polyv);
}

/* the tables of the plain call (mrefinelev=0: offsets -6..6) */
void oracle_loctab_level(double eps, int ipoly, int n, double delta, double boxdim, double *tab,
                         double *tabx, double *tabxx, int *ind) {
    oracle_loctab_level_m(eps, ipoly, n, delta, boxdim, 0, tab, tabx, tabxx, ind);
}

/* ------------------------------------------------------------------ */
/* tree lists (common/tree_routs_nd.f)                                  */
/* ------------------------------------------------------------------ */

typedef struct {
    int nlevels, nboxes;
    int *laddr, *ilevel, *iparent, *nchild, *ichild, *ncoll, *coll;
    const double *centers, *boxsize;
} tree_t;

static void tree_view(tree_t *t, int nlevels, int *itree, const int *iptr, const double *centers,
                      const double *boxsize) {
    t->nlevels = nlevels;
    t->laddr = itree + iptr[0] - 1;
    t->ilevel = itree + iptr[1] - 1;
    t->iparent = itree + iptr[2] - 1;
    t->nchild = itree + iptr[3] - 1;
    t->ichild = itree + iptr[4] - 1;
    t->ncoll = itree + iptr[5] - 1;
    t->coll = itree + iptr[6] - 1;
    t->nboxes = itree[2 * nlevels + 1];
    t->centers = centers;
    t->boxsize = boxsize;
}
#define LFIRST(t, l) ((t)->laddr[2 * (l)] - 1)
#define LLAST(t, l) ((t)->laddr[2 * (l) + 1] - 1)

static double per_dist(double d, int iperiod, double bs0) {
    d = fabs(d);
    if (iperiod == 1) {
        double dp = bs0 - d;
        if (dp < d) d = dp;
    }
    return d;
}

/* computecoll (common/tree_routs_nd.f:106-242): rewrites ncoll/coll. */
void oracle_computecoll(tree_t *t, int iperiod) {
    double bs0 = t->boxsize[0];
    for (int i = 0; i < t->nboxes; i++) {
        t->ncoll[i] = 0;
        for (int j = 0; j < 9; j++) t->coll[9 * i + j] = -1;
    }
    t->ncoll[0] = 1;
    t->coll[0] = 1;
    for (int l = 1; l <= t->nlevels; l++) {
        for (int b = LFIRST(t, l); b <= LLAST(t, l); b++) {
            int dad = t->iparent[b] - 1;
            for (int i = 0; i < t->ncoll[dad]; i++) {
                int jb = t->coll[9 * dad + i] - 1;
                for (int c = 0; c < 4; c++) {
                    int kb = t->ichild[4 * jb + c];
                    if (kb <= 0) continue;
                    kb -= 1;
                    int nb = 1;
                    for (int d = 0; d < 2; d++) {
                        double dis = per_dist(t->centers[2 * kb + d] - t->centers[2 * b + d], iperiod, bs0);
                        if (dis > 1.05 * t->boxsize[l]) { nb = 0; break; }
                    }
                    if (nb) t->coll[9 * b + t->ncoll[b]++] = kb + 1;
                }
            }
        }
    }
}

/* gnd_find_pwexp_boxes (common/tree_routs_nd.f:1266-1368). */
static void find_pwexp_boxes(const tree_t *t, int npwlevel, int *ifpwexp) {
    for (int i = 0; i < t->nboxes; i++) ifpwexp[i] = 0;
    if (npwlevel <= 0) ifpwexp[0] = 1;
    int nc = npwlevel <= 0 ? 0 : npwlevel;
    if (npwlevel > t->nlevels) return;
    for (int b = LFIRST(t, nc); b <= LLAST(t, nc); b++) {
        if (t->nchild[b] == 0) {
            for (int j = 0; j < t->ncoll[b]; j++) {
                int jb = t->coll[9 * b + j] - 1;
                if (t->nchild[jb] > 0) { ifpwexp[b] = 1; break; }
            }
        } else ifpwexp[b] = 1;
    }
    for (int l = nc + 1; l <= t->nlevels; l++)
        for (int b = LFIRST(t, l); b <= LLAST(t, l); b++) ifpwexp[b] = 1;
}

/* gnd_compute_modified_list1 (common/tree_routs_nd.f:864-1055).  list1 is
 * stored with stride mnl (0-based box ids). */
static void modified_list1(const tree_t *t, int npwlevel, const int *ifpwexp, int iperiod, int mnl,
                           int *nlist1, int *list1) {
    double bs0 = t->boxsize[0];
    for (int i = 0; i < t->nboxes; i++) nlist1[i] = 0;
    if (t->nchild[0] == 0) { nlist1[0] = 1; list1[0] = 0; }
    int nlevend = t->nlevels;
    if (npwlevel < t->nlevels) nlevend = npwlevel;
    for (int l = 1; l <= nlevend; l++) {
        for (int b = LFIRST(t, l); b <= LLAST(t, l); b++) {
            if (t->nchild[b] != 0) continue;
            int dad = t->iparent[b] - 1;
            for (int i = 0; i < t->ncoll[b]; i++) {
                int jb = t->coll[9 * b + i] - 1;
                if (t->nchild[jb] == 0) {
                    if (ifpwexp[b] != 1 || jb != b) list1[mnl * b + nlist1[b]++] = jb;
                } else if (l < npwlevel) {
                    double dt = 1.05 * (t->boxsize[l] + t->boxsize[l + 1]) / 2.0;
                    for (int c = 0; c < 4; c++) {
                        int kb = t->ichild[4 * jb + c] - 1;
                        if (t->nchild[kb] != 0) continue;
                        int ok = 1;
                        for (int d = 0; d < 2; d++) {
                            double dis = per_dist(t->centers[2 * kb + d] - t->centers[2 * b + d], iperiod, bs0);
                            if (dis >= dt) { ok = 0; break; }
                        }
                        if (ok) list1[mnl * b + nlist1[b]++] = kb;
                    }
                }
            }
            for (int i = 0; i < t->ncoll[dad]; i++) {
                int jb = t->coll[9 * dad + i] - 1;
                if (t->nchild[jb] != 0) continue;
                double dt = 1.05 * (t->boxsize[l] + t->boxsize[l - 1]) / 2.0;
                int ok = 1;
                for (int d = 0; d < 2; d++) {
                    double dis = per_dist(t->centers[2 * jb + d] - t->centers[2 * b + d], iperiod, bs0);
                    if (dis >= dt) { ok = 0; break; }
                }
                if (ok) list1[mnl * b + nlist1[b]++] = jb;
            }
        }
    }
}

/* gnd_compute_listpw (common/tree_routs_nd.f:1564-1630), stride 8. */
static void compute_listpw(const tree_t *t, int npwlevel, int *nlistpw, int *listpw) {
    for (int i = 0; i < t->nboxes; i++) nlistpw[i] = 0;
    if (npwlevel > t->nlevels) return;
    int nc = npwlevel < 0 ? 0 : npwlevel;
    for (int b = LFIRST(t, nc); b <= LLAST(t, nc); b++)
        for (int i = 0; i < t->ncoll[b]; i++) {
            int jb = t->coll[9 * b + i] - 1;
            if (jb != b && (t->nchild[jb] > 0 || t->nchild[b] > 0)) listpw[8 * b + nlistpw[b]++] = jb;
        }
}

/* ------------------------------------------------------------------ */
/* plane-wave translation operators (bfgt/bfgt_pwrouts.f)               */
/* ------------------------------------------------------------------ */

/* merge/split diagonals for klev=1..nmax, quadrants pp,mp,pm,mm;
 * w[(klev-1)*4*nexp + q*nexp + j], j = j2 + npw*j1 (x index j2 fastest)
 * (g2d_mk_merge_split_pw_matrices, bfgt/bfgt_pwrouts.f:1645-1703). */
static void mk_merge_split(double xmin, int npw, const double *ts, int nmax, cplx *w) {
    int H = (npw + 1) / 2, nexp = npw * H;
    cplx *ww = (cplx *)malloc(sizeof(cplx) * npw * (nmax > 0 ? nmax : 1));
    for (int j = 0; j < npw; j++) {
        cplx z = cexp(I * ts[j] * xmin);
        for (int k = 0; k < nmax; k++) {
            ww[j + npw * k] = z;
            z = z * z;
        }
    }
    for (int k = 0; k < nmax; k++)
        for (int j1 = 0; j1 < H; j1++)
            for (int j2 = 0; j2 < npw; j2++) {
                int j = j2 + npw * j1;
                cplx a = ww[j2 + npw * k], b = ww[j1 + npw * k];
                w[(size_t)k * 4 * nexp + 0 * nexp + j] = a * b;
                w[(size_t)k * 4 * nexp + 1 * nexp + j] = conj(a) * b;
                w[(size_t)k * 4 * nexp + 2 * nexp + j] = a * conj(b);
                w[(size_t)k * 4 * nexp + 3 * nexp + j] = conj(a) * conj(b);
            }
    free(ww);
}

/* M2L diagonals for offsets (kx,ky) in {-1,0,1}^2, index (ky+1)*3+(kx+1)
 * (g2d_mk_pw_translation_matrices, bfgt/bfgt_pwrouts.f:1415-1471). */
static void mk_translation(double xmin, int npw, const double *ts, cplx *w) {
    int H = (npw + 1) / 2, nexp = npw * H;
    cplx *ww = (cplx *)malloc(sizeof(cplx) * npw * 3); /* offsets -1,0,1 */
    for (int j = 0; j < npw; j++) {
        cplx z = cexp(I * ts[j] * xmin);
        ww[j + npw * 1] = 1.0;
        ww[j + npw * 2] = z;
        ww[j + npw * 0] = conj(z);
    }
    for (int ky = -1; ky <= 1; ky++)
        for (int kx = -1; kx <= 1; kx++) {
            int k = (ky + 1) * 3 + (kx + 1);
            for (int j1 = 0; j1 < H; j1++)
                for (int j2 = 0; j2 < npw; j2++)
                    w[(size_t)k * nexp + j2 + npw * j1] = ww[j2 + npw * (kx + 1)] * ww[j1 + npw * (ky + 1)];
        }
    free(ww);
}

/* gnd_find_msshift_ind (bfgt/bfgt_pwrouts.f:1946-1976), 0-based. */
static int msshift_ind(const double *tc, const double *sc) {
    int k = 0;
    if (tc[0] - sc[0] < 0) k += 1;
    if (tc[1] - sc[1] < 0) k += 2;
    return k;
}

/* gnd_find_pwshift_ind (bfgt/bfgt_pwrouts.f:1848-1892), 0-based. */
static int pwshift_ind(int iperiod, const double *tc, const double *sc, double bs0, double xmin) {
    int jx[2];
    for (int i = 0; i < 2; i++) {
        int j = (int)lround((tc[i] - sc[i]) / xmin);
        if (iperiod == 1) {
            int jp = (int)lround((tc[i] - sc[i] - bs0) / xmin);
            int jm = (int)lround((tc[i] - sc[i] + bs0) / xmin);
            if (abs(j) > abs(jp)) j = jp;
            if (abs(j) > abs(jm)) j = jm;
        }
        jx[i] = j + 1;
    }
    return jx[0] + 3 * jx[1];
}

/* gnd_find_loctab_ind (bfgt/bfgt_local.f:1729-1765): truncation of
 * (t-s)/(b_src/2^(2+mrefinelev)) toward zero, nearest periodic image. */
static void loctab_ind(int iperiod, const double *tc, const double *sc, double sbs, double bs0, int mref,
                       int *ixy) {
    double bs = sbs / (double)(1 << (2 + mref));
    for (int i = 0; i < 2; i++) {
        double dx = (tc[i] - sc[i]) / bs;
        if (iperiod == 1) {
            double dp = dx - bs0 / bs, dm = dx + bs0 / bs;
            if (fabs(dx) > fabs(dp)) dx = dp;
            if (fabs(dx) > fabs(dm)) dx = dm;
        }
        ixy[i] = (int)dx;
    }
}

/* ------------------------------------------------------------------ */
/* leaf kernels                                                         */
/* ------------------------------------------------------------------ */

/* g2d_tens_prod_to_pw (bfgt/bfgt_pwrouts.f:98-147): pw(k1,k2,d). */
static void tens_prod_to_pw(int nd, int n, const double *f, int npw, const cplx *tab, cplx *pw) {
    int H = (npw + 1) / 2;
    cplx ff[MAXN * 100];
    for (int d = 0; d < nd; d++) {
        for (int m1 = 0; m1 < n; m1++)
            for (int k2 = 0; k2 < H; k2++) {
                cplx cd = 0;
                for (int m2 = 0; m2 < n; m2++) cd += tab[m2 + n * k2] * f[(m1 + n * m2) * nd + d];
                ff[m1 + n * k2] = cd;
            }
        for (int k2 = 0; k2 < H; k2++)
            for (int k1 = 0; k1 < npw; k1++) {
                cplx cd = 0;
                for (int m1 = 0; m1 < n; m1++) cd += tab[m1 + n * k1] * ff[m1 + n * k2];
                pw[k1 + npw * k2 + (size_t)npw * H * d] = cd;
            }
    }
}

/* g2d_pw2pot / g2d_pw2pg / g2d_pw2pgh (bfgt/bfgt_pwrouts.f:549-784). */
static void pw2pgh(int nd, int n, int npw, const cplx *pw, const cplx *tp, const cplx *tx, const cplx *txx,
                   int ifpgh, double *pot, double *grad, double *hess) {
    int H = (npw + 1) / 2, npw2 = npw / 2;
    cplx ff[100 * MAXN], ffx[100 * MAXN], ffxx[100 * MAXN];
    for (int d = 0; d < nd; d++) {
        const cplx *p = pw + (size_t)npw * H * d;
        for (int m2 = 0; m2 < H; m2++)
            for (int k1 = 0; k1 < n; k1++) {
                cplx c0 = 0, c1 = 0, c2 = 0;
                for (int m1 = 0; m1 < npw; m1++) {
                    cplx v = p[m1 + npw * m2];
                    c0 += tp[m1 + npw * k1] * v;
                    if (ifpgh >= 2) c1 += tx[m1 + npw * k1] * v;
                    if (ifpgh >= 3) c2 += txx[m1 + npw * k1] * v;
                }
                ff[m2 + H * k1] = c0;
                ffx[m2 + H * k1] = c1;
                ffxx[m2 + H * k1] = c2;
            }
        for (int k2 = 0; k2 < n; k2++)
            for (int k1 = 0; k1 < n; k1++) {
                cplx cd = 0, cx = 0, cy = 0, cxx = 0, cxy = 0, cyy = 0;
                for (int m2 = 0; m2 < npw2; m2++) {
                    cplx a = ff[m2 + H * k1];
                    cd += tp[m2 + npw * k2] * a;
                    if (ifpgh >= 2) {
                        cy += tx[m2 + npw * k2] * a;
                        cx += tp[m2 + npw * k2] * ffx[m2 + H * k1];
                    }
                    if (ifpgh >= 3) {
                        cyy += txx[m2 + npw * k2] * a;
                        cxy += tx[m2 + npw * k2] * ffx[m2 + H * k1];
                        cxx += tp[m2 + npw * k2] * ffxx[m2 + H * k1];
                    }
                }
                if (H > npw2) {
                    int m2 = H - 1;
                    cplx a = ff[m2 + H * k1];
                    cd += tp[m2 + npw * k2] * a / 2.0;
                    if (ifpgh >= 2) {
                        cy += tx[m2 + npw * k2] * a / 2.0;
                        cx += tp[m2 + npw * k2] * ffx[m2 + H * k1] / 2.0;
                    }
                    if (ifpgh >= 3) {
                        cyy += txx[m2 + npw * k2] * a / 2.0;
                        cxy += tx[m2 + npw * k2] * ffx[m2 + H * k1] / 2.0;
                        cxx += tp[m2 + npw * k2] * ffxx[m2 + H * k1] / 2.0;
                    }
                }
                int q = k1 + n * k2;
                pot[q * nd + d] += creal(cd) * 2;
                if (ifpgh >= 2) {
                    grad[(q * 2 + 0) * nd + d] += creal(cx) * 2;
                    grad[(q * 2 + 1) * nd + d] += creal(cy) * 2;
                }
                if (ifpgh >= 3) {
                    hess[(q * 3 + 0) * nd + d] += creal(cxx) * 2;
                    hess[(q * 3 + 1) * nd + d] += creal(cxy) * 2;
                    hess[(q * 3 + 2) * nd + d] += creal(cyy) * 2;
                }
            }
    }
}

/* g2d_tens_prod_to_potloc/pgloc/pghloc (bfgt/bfgt_local.f:315-545) for
 * table offsets (ix,iy) of one target level. */
static void tens_prod_to_pghloc(int nd, int n, const double *f, int ifpgh, double *pot, double *grad,
                                double *hess, const double *tab, const double *tabx, const double *tabxx,
                                const int *ind, int nloctab, const int *ixy) {
    size_t nn = (size_t)n * n;
    int qx = ixy[0] + nloctab, qy = ixy[1] + nloctab;
    const int *ix = ind + qx * 2 * (n + 1), *iy = ind + qy * 2 * (n + 1);
    const double *Tx = tab + qx * nn, *Txd = tabx + qx * nn, *Txdd = tabxx + qx * nn;
    const double *Ty = tab + qy * nn, *Tyd = tabx + qy * nn, *Tydd = tabxx + qy * nn;
    int kx0 = ix[2 * n] - 1, kx1 = ix[2 * n + 1] - 1;
    int ky0 = iy[2 * n] - 1, ky1 = iy[2 * n + 1] - 1;
    if (kx1 - kx0 + 1 == 0 || ky1 - ky0 + 1 == 0) return;
    double ff[MAXN * MAXN], ffx[MAXN * MAXN], ffxx[MAXN * MAXN];
    for (int d = 0; d < nd; d++) {
        for (int j2 = 0; j2 < n; j2++)
            for (int k1 = kx0; k1 <= kx1; k1++) {
                double c0 = 0, c1 = 0, c2 = 0;
                for (int j1 = ix[2 * k1] - 1; j1 <= ix[2 * k1 + 1] - 1; j1++) {
                    double v = f[(j1 + n * j2) * nd + d];
                    c0 += Tx[j1 + n * k1] * v;
                    if (ifpgh >= 2) c1 += Txd[j1 + n * k1] * v;
                    if (ifpgh >= 3) c2 += Txdd[j1 + n * k1] * v;
                }
                ff[k1 + n * j2] = c0;
                ffx[k1 + n * j2] = c1;
                ffxx[k1 + n * j2] = c2;
            }
        for (int k2 = ky0; k2 <= ky1; k2++)
            for (int k1 = kx0; k1 <= kx1; k1++) {
                double cd = 0, cx = 0, cy = 0, cxx = 0, cxy = 0, cyy = 0;
                for (int j2 = iy[2 * k2] - 1; j2 <= iy[2 * k2 + 1] - 1; j2++) {
                    double a = ff[k1 + n * j2];
                    cd += Ty[j2 + n * k2] * a;
                    if (ifpgh >= 2) {
                        cy += Tyd[j2 + n * k2] * a;
                        cx += Ty[j2 + n * k2] * ffx[k1 + n * j2];
                    }
                    if (ifpgh >= 3) {
                        cyy += Tydd[j2 + n * k2] * a;
                        cxy += Tyd[j2 + n * k2] * ffx[k1 + n * j2];
                        cxx += Ty[j2 + n * k2] * ffxx[k1 + n * j2];
                    }
                }
                int q = k1 + n * k2;
                pot[q * nd + d] += cd;
                if (ifpgh >= 2) {
                    grad[(q * 2 + 0) * nd + d] += cx;
                    grad[(q * 2 + 1) * nd + d] += cy;
                }
                if (ifpgh >= 3) {
                    hess[(q * 3 + 0) * nd + d] += cxx;
                    hess[(q * 3 + 1) * nd + d] += cxy;
                    hess[(q * 3 + 2) * nd + d] += cyy;
                }
            }
    }
}

/* ------------------------------------------------------------------ */
/* driver                                                               */
/* ------------------------------------------------------------------ */

/* state handed from the transform to the refinement steps */
typedef struct newtree_in {
    int nd, n, ipoly, ifpgh, iperiod, npw, npwlevel, nlevstart, mref, nbmax;
    int *nlevels;
    tree_t *T;
    double *centers;
    const double *boxsize, *fvals;
    double *pot, *grad, *hess;
    const double *xq, *wq, *umat;
    double eps;
    cplx *exps;
    const int64_t *aloc;
    size_t blk;
    const cplx *wms, *tp, *tx, *txx;
    const int *ifpwexp, *nlist1, *list1;
    int mnl;
    const double *tab, *tabx, *tabxx;
    const int *ind;
    double *timeinfo;
    int *info;
} newtree_in;
static int newtree_steps(const newtree_in *S);

/*
 * oracle_boxfgt_nt: boxfgt + bfgtmain (bfgt/boxfgt.f:5-286, 291-1115,
 * 1172-1483) for ndim=2 and ifpghtarg=0.  Same side effects as the
 * reference: itree's colleague lists are recomputed, pot/grad/hess are
 * accumulated (only leaf boxes are written), timeinfo(1..20) is zeroed and
 * slots 1-7 (9-11 with ifnewtree=1) are filled.  With ifnewtree=1 the tree
 * (itree, *nlevels, centers) and pot/grad/hess are then refined, coarsened
 * and 2:1-rebalanced to resolve pot (STEPS 8-10); nboxes is the capacity of
 * the box arrays.  info (may be NULL) receives {npwlevel, npw, nlevstart,
 * nlevend, nleafbox, ndirect_pairs, boxes created by STEP 8, boxes whose
 * refinement stopped at the finest level}.  Returns 0, or a nonzero code
 * for unsupported requests (7: the refined tree outgrows nboxes).
 */
int oracle_boxfgt_nt(int nd, double delta, double eps, int ipoly, int iperiod, int norder, int npbox,
                     int nboxes, int *nlevels_io, int *itree, const int *iptr, double *centers_io,
                     const double *boxsize, const double *fvals, int ifpgh, double *pot, double *grad,
                     double *hess, int ifnewtree, double *timeinfo, int *info) {
    const int nlevels = *nlevels_io;
    const double *centers = centers_io;
    if (norder < 1 || norder > MAXN || npbox != norder * norder) return 3;
    if (ipoly != 0 && ipoly != 1) return 4;
    if (ifpgh < 1 || ifpgh > 3) return 5;
    int n = norder;
    tree_t T;
    tree_view(&T, nlevels, itree, iptr, centers, boxsize);
    int nb = T.nboxes;

    /* cutoff level and PW length, boxfgt.f:133-167 */
    double dcutoff = sqrt(delta * log(1.0 / eps));
    int npwlevel = nlevels + 1;
    for (int i = nlevels; i >= 0; i--)
        if (boxsize[i] >= dcutoff) { npwlevel = i; break; }
    if (boxsize[nlevels] > dcutoff) npwlevel = nlevels + 1;
    if (boxsize[0] <= dcutoff) npwlevel = 0;
    if (npwlevel <= 1 && iperiod == 1) npwlevel = 0;
    int npw = 0;
    double pmax = 0;
    if (npwlevel <= nlevels) {
        if (npwlevel <= 0) oracle_fgtpwterms(boxsize[0], delta, eps, iperiod, &pmax, &npw);
        else oracle_fgtpwterms(2 * boxsize[npwlevel], delta, eps, 0, &pmax, &npw);
    }
    if (npw > 200) return 6; /* reference's fixed ws(200), ts(200) */

    for (int i = 0; i < 20; i++) timeinfo[i] = 0;
    double t1 = wtime();

    /* STEP 0 (bfgtmain :489-698) */
    double bs0 = boxsize[0];
    int iperiod0 = iperiod;
    double delta0 = delta;
    double ws[200], ts[200];
    if (npwlevel <= 1 && iperiod == 1) {
        oracle_periodic_pwnodes(bs0, delta, npw, ws, ts);
        iperiod = 0;
        npwlevel = 0;
        delta0 = 1.0;
        oracle_computecoll(&T, iperiod);
    } else {
        oracle_computecoll(&T, iperiod);
        if (npw > 0) oracle_pwnodes(pmax, npw, ws, ts);
    }
    double xq[MAXN], wq[MAXN], umat[MAXN * MAXN];
    if (ipoly == 0) oracle_legeexps(n, xq, wq, umat);
    else oracle_chebexps(n, xq, wq, umat);

    int nlevend = nlevels;
    if (npwlevel < nlevels) nlevend = npwlevel;
    if (npwlevel == 0 && iperiod0 == 1) nlevend = -1;
    int nleafbox = 0;
    for (int l = 1; l <= nlevels; l++)
        for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++)
            if (T.nchild[b] == 0) nleafbox++;

    int *ifpwexp = (int *)malloc(sizeof(int) * nb);
    find_pwexp_boxes(&T, npwlevel, ifpwexp);
    const int mnl = 32;
    int *nlist1 = (int *)malloc(sizeof(int) * nb);
    int *list1 = (int *)malloc(sizeof(int) * nb * mnl);
    modified_list1(&T, npwlevel, ifpwexp, iperiod, mnl, nlist1, list1);

    /* near-field tables: with ifnewtree=1 every target may be refined
     * mrefinelev=3 times (bfgt/boxfgt.f:619-636) */
    const int mref = ifnewtree == 1 ? 3 : 0;
    const int nloc = loctab_half(mref), nslot = 2 * nloc + 1;
    size_t nn = (size_t)n * n;
    double *tab = (double *)malloc(sizeof(double) * nn * nslot * (nlevels + 1));
    double *tabx = (double *)malloc(sizeof(double) * nn * nslot * (nlevels + 1));
    double *tabxx = (double *)malloc(sizeof(double) * nn * nslot * (nlevels + 1));
    int *ind = (int *)malloc(sizeof(int) * 2 * (n + 1) * nslot * (nlevels + 1));
    for (int l = 0; l <= nlevels; l++)
        oracle_loctab_level_m(eps, ipoly, n, delta, boxsize[l], mref, tab + nn * nslot * l,
                              tabx + nn * nslot * l, tabxx + nn * nslot * l, ind + 2 * (n + 1) * nslot * l);

    cplx *exps = NULL;
    double t2;
    /* kept for the refinement step (STEP 8 shifts and evaluates the local
     * expansions of refined leaves) */
    cplx *tpoly = NULL, *tp = NULL, *tx = NULL, *txx = NULL, *wms = NULL, *wsh = NULL;
    int64_t *amp = NULL, *aloc = NULL;
    int *nlistpw = NULL, *listpw = NULL;
    int has_pw = 0;
    if (!(npwlevel >= nlevels && npwlevel > 0)) {
        has_pw = 1;
        int nlevstart = npwlevel > 0 ? npwlevel : 0;
        int H = (npw + 1) / 2, nexp = npw * H;
        tpoly = (cplx *)malloc(sizeof(cplx) * n * npw * (nlevels + 1));
        tp = (cplx *)malloc(sizeof(cplx) * n * npw * (nlevels + 1));
        tx = (cplx *)malloc(sizeof(cplx) * n * npw * (nlevels + 1));
        txx = (cplx *)malloc(sizeof(cplx) * n * npw * (nlevels + 1));
        for (int l = nlevstart; l <= nlevels; l++) {
            oracle_poly2pw(n, ipoly, npw, ws, ts, delta0, boxsize[l], tpoly + (size_t)n * npw * l);
            oracle_pw2pgh(n, npw, ts, xq, delta0, boxsize[l], tp + (size_t)n * npw * l,
                          tx + (size_t)n * npw * l, txx + (size_t)n * npw * l);
        }
        int nmax = nlevels - (npwlevel > 0 ? npwlevel : 0);
        wms = (cplx *)malloc(sizeof(cplx) * nexp * 4 * (nmax > 0 ? nmax : 1));
        mk_merge_split(boxsize[nlevels] / sqrt(delta0) / 2, npw, ts, nmax, wms);
        wsh = (cplx *)malloc(sizeof(cplx) * nexp * 9);
        mk_translation(boxsize[nlevstart] / sqrt(delta0), npw, ts, wsh);
        double xmin = boxsize[nlevstart];
        nlistpw = (int *)malloc(sizeof(int) * nb);
        listpw = (int *)malloc(sizeof(int) * nb * 8);
        compute_listpw(&T, npwlevel, nlistpw, listpw);

        /* expansion storage (gnd_mpalloc, boxfgt.f:1598-1702): mp for
         * every box at levels >= nlevstart, loc separately at nlevstart,
         * loc aliased to mp below it */
        size_t blk = (size_t)nd * nexp;
        amp = (int64_t *)malloc(sizeof(int64_t) * nb);
        aloc = (int64_t *)malloc(sizeof(int64_t) * nb);
        int64_t off = 0;
        for (int l = nlevstart; l <= nlevels; l++)
            for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++) amp[b] = off, off += blk;
        for (int b = LFIRST(&T, nlevstart); b <= LLAST(&T, nlevstart); b++) aloc[b] = off, off += blk;
        for (int l = nlevstart + 1; l <= nlevels; l++)
            for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++) aloc[b] = amp[b];
        exps = (cplx *)calloc((size_t)off, sizeof(cplx));
        t2 = wtime();
        timeinfo[0] += t2 - t1;

        /* STEP 1: leaf values -> mp (boxfgt.f:728-744) */
        t1 = wtime();
        for (int l = nlevels; l >= nlevstart; l--)
            for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++)
                if (ifpwexp[b] == 1 && T.nchild[b] == 0)
                    tens_prod_to_pw(nd, n, fvals + (size_t)b * npbox * nd, npw, tpoly + (size_t)n * npw * l,
                                    exps + amp[b]);
        t2 = wtime();
        timeinfo[1] += t2 - t1;

        /* STEP 2: merge mp children -> parent (boxfgt.f:764-783) */
        t1 = wtime();
        for (int l = nlevels - 1; l >= nlevstart; l--) {
            int klev = nlevels - l;
            for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++) {
                int nc = T.nchild[b];
                if (nc > 0) memset(exps + amp[b], 0, sizeof(cplx) * blk);
                for (int c = 0; c < nc; c++) {
                    int jb = T.ichild[4 * b + c] - 1;
                    int k = msshift_ind(centers + 2 * b, centers + 2 * jb);
                    const cplx *w = wms + (size_t)(klev - 1) * 4 * nexp + (size_t)k * nexp;
                    for (int d = 0; d < nd; d++)
                        for (int j = 0; j < nexp; j++)
                            exps[amp[b] + (size_t)d * nexp + j] += exps[amp[jb] + (size_t)d * nexp + j] * w[j];
                }
            }
        }
        t2 = wtime();
        timeinfo[2] += t2 - t1;

        /* STEP 3: mp -> loc at the cutoff level (boxfgt.f:805-822) */
        t1 = wtime();
        for (int b = LFIRST(&T, nlevstart); b <= LLAST(&T, nlevstart); b++) {
            memcpy(exps + aloc[b], exps + amp[b], sizeof(cplx) * blk);
            for (int j = 0; j < nlistpw[b]; j++) {
                int jb = listpw[8 * b + j];
                int k = pwshift_ind(iperiod, centers + 2 * b, centers + 2 * jb, bs0, xmin);
                const cplx *w = wsh + (size_t)k * nexp;
                for (int d = 0; d < nd; d++)
                    for (int q = 0; q < nexp; q++)
                        exps[aloc[b] + (size_t)d * nexp + q] += exps[amp[jb] + (size_t)d * nexp + q] * w[q];
            }
        }
        t2 = wtime();
        timeinfo[3] += t2 - t1;

        /* STEP 4: split loc parent -> children (boxfgt.f:845-863) */
        t1 = wtime();
        for (int l = nlevstart; l <= nlevels - 1; l++) {
            int klev = nlevels - l;
            for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++) {
                for (int c = 0; c < T.nchild[b]; c++) {
                    int jb = T.ichild[4 * b + c] - 1;
                    int k = msshift_ind(centers + 2 * jb, centers + 2 * b);
                    const cplx *w = wms + (size_t)(klev - 1) * 4 * nexp + (size_t)k * nexp;
                    for (int d = 0; d < nd; d++)
                        for (int j = 0; j < nexp; j++)
                            exps[aloc[jb] + (size_t)d * nexp + j] = exps[aloc[b] + (size_t)d * nexp + j] * w[j];
                }
            }
        }
        t2 = wtime();
        timeinfo[4] += t2 - t1;

        /* STEP 5: evaluate loc on PW leaves (boxfgt.f:884-900) */
        t1 = wtime();
        for (int l = nlevstart; l <= nlevels; l++)
            for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++)
                if (ifpwexp[b] == 1 && T.nchild[b] == 0)
                    pw2pgh(nd, n, npw, exps + aloc[b], tp + (size_t)n * npw * l, tx + (size_t)n * npw * l,
                           txx + (size_t)n * npw * l, ifpgh, pot + (size_t)b * npbox * nd,
                           grad + (size_t)b * npbox * nd * 2, hess + (size_t)b * npbox * nd * 3);
        t2 = wtime();
        timeinfo[5] += t2 - t1;
        if (info) { info[2] = nlevstart; }
    } else {
        t2 = wtime();
        timeinfo[0] += t2 - t1;
        if (info) { info[2] = -1; }
    }

    /* STEP 6: direct interactions (boxfgt.f:923-1071) */
    t1 = wtime();
    long npairs = 0;
    if (!(npwlevel == 0 && iperiod == 0)) {
        for (int l = 0; l <= nlevend; l++) {
            const double *Tl = tab + nn * nslot * l, *Txl = tabx + nn * nslot * l,
                         *Txxl = tabxx + nn * nslot * l;
            const int *Il = ind + 2 * (n + 1) * nslot * l;
            for (int b = LFIRST(&T, l); b <= LLAST(&T, l); b++) {
                double *pb = pot + (size_t)b * npbox * nd;
                double *gb = grad ? grad + (size_t)b * npbox * nd * 2 : NULL;
                double *hb = hess ? hess + (size_t)b * npbox * nd * 3 : NULL;
                for (int j = 0; j < nlist1[b]; j++) {
                    int jb = list1[mnl * b + j];
                    int jl = T.ilevel[jb];
                    double bs = boxsize[jl];
                    const double *fj = fvals + (size_t)jb * npbox * nd;
                    int ixy[2];
                    if (iperiod == 1 && jl == 1) {
                        loctab_ind(0, centers + 2 * b, centers + 2 * jb, bs, bs0, mref, ixy);
                        tens_prod_to_pghloc(nd, n, fj, ifpgh, pb, gb, hb, Tl, Txl, Txxl, Il, nloc, ixy);
                        npairs++;
                        /* periodic images of a level-1 source (boxfgt.f:951-1053) */
                        int im[2], nim = 0;
                        double dxy[2], ct[3][2], bsh = bs * 0.25;
                        for (int i = 0; i < 2; i++) {
                            im[i] = 0;
                            dxy[i] = centers[2 * jb + i] - centers[2 * b + i];
                            if (fabs(dxy[i]) > bsh) im[i] = 1;
                            nim += im[i];
                        }
                        if (nim > 0) {
                            for (int i = 0; i < 2; i++) {
                                if (dxy[i] < -bsh) ct[0][i] = centers[2 * jb + i] + bs0;
                                else if (dxy[i] > bsh) ct[0][i] = centers[2 * jb + i] - bs0;
                                else ct[0][i] = centers[2 * jb + i];
                            }
                            if (nim == 2) {
                                ct[1][0] = centers[2 * jb + 0];
                                ct[1][1] = ct[0][1];
                                ct[2][0] = ct[0][0];
                                ct[2][1] = centers[2 * jb + 1];
                            }
                            for (int k = 0; k < (1 << nim) - 1; k++) {
                                loctab_ind(0, centers + 2 * b, ct[k], bs, bs0, mref, ixy);
                                tens_prod_to_pghloc(nd, n, fj, ifpgh, pb, gb, hb, Tl, Txl, Txxl, Il, nloc,
                                                    ixy);
                                npairs++;
                            }
                        }
                    } else {
                        loctab_ind(iperiod, centers + 2 * b, centers + 2 * jb, bs, bs0, mref, ixy);
                        tens_prod_to_pghloc(nd, n, fj, ifpgh, pb, gb, hb, Tl, Txl, Txxl, Il, nloc, ixy);
                        npairs++;
                    }
                }
            }
        }
    }
    t2 = wtime();
    timeinfo[6] += t2 - t1;
    if (info) {
        info[0] = npwlevel;
        info[1] = npw;
        info[3] = nlevend;
        info[4] = nleafbox;
        info[5] = (int)npairs;
        info[6] = info[7] = 0;
    }
    int rc = 0;
    if (ifnewtree == 1) {
        newtree_in S = {nd, n, ipoly, ifpgh, iperiod, npw, npwlevel, has_pw ? (npwlevel > 0 ? npwlevel : 0) : -1,
                        mref, nboxes, nlevels_io, &T, centers_io, boxsize, fvals, pot, grad, hess,
                        xq, wq, umat, eps, exps, aloc, (size_t)nd * npw * ((npw + 1) / 2), wms, tp, tx, txx,
                        ifpwexp, nlist1, list1, mnl, tab, tabx, tabxx, ind, timeinfo, info};
        rc = newtree_steps(&S);
    }
    free(exps);
    free(tpoly); free(tp); free(tx); free(txx); free(wms); free(wsh);
    free(nlistpw); free(listpw); free(amp); free(aloc);
    free(ifpwexp); free(nlist1); free(list1);
    free(tab); free(tabx); free(tabxx); free(ind);
    return rc;
}

/* the plain call (ifnewtree=0): the tree is only read (colleague lists
 * rewritten) */
int oracle_boxfgt(int nd, double delta, double eps, int ipoly, int iperiod, int norder, int npbox,
                  int nboxes, int nlevels, int *itree, const int *iptr, const double *centers,
                  const double *boxsize, const double *fvals, int ifpgh, double *pot, double *grad,
                  double *hess, double *timeinfo, int *info) {
    int nl = nlevels;
    return oracle_boxfgt_nt(nd, delta, eps, ipoly, iperiod, norder, npbox, nboxes, &nl, itree, iptr,
                            (double *)centers, boxsize, fvals, ifpgh, pot, grad, hess, 0, timeinfo, info);
}

/* ------------------------------------------------------------------ */
/* ifnewtree=1: refine, coarsen and rebalance the tree to resolve pot    */
/* (bfgtmain STEPS 8-10, bfgt/boxfgt.f:1172-1483)                        */
/* ------------------------------------------------------------------ */


/* ortho_trans_2d with itype=0 (common/tensor_prod_routs_nd.f:21-130): the
 * tensor-product coefficients c[(k1 + n k2) nf + f] of nf functions per
 * point, u[k + n j] = U(coefficient k, value j). */
static void nt_trans(int nf, int n, const double *u, const double *v, double *c) {
    double w[MAXN * MAXN];
    for (int f = 0; f < nf; f++) {
        for (int j2 = 0; j2 < n; j2++)
            for (int k1 = 0; k1 < n; k1++) {
                double s = 0;
                for (int j1 = 0; j1 < n; j1++) s += u[k1 + n * j1] * v[(j1 + n * j2) * nf + f];
                w[k1 + n * j2] = s;
            }
        for (int k2 = 0; k2 < n; k2++)
            for (int k1 = 0; k1 < n; k1++) {
                double s = 0;
                for (int j2 = 0; j2 < n; j2++) s += u[k2 + n * j2] * w[k1 + n * j2];
                c[(k1 + n * k2) * nf + f] = s;
            }
    }
}

/* ortho_eval_2d (common/tensor_prod_routs_nd.f:527-562): the expansion at
 * m x m points, px[k + n a] = p_k(x_a), py likewise; out[(a + m b) nf + f]. */
static void nt_eval(int nf, int n, const double *c, int m, const double *px, const double *py, double *out) {
    double w[MAXN * MAXN];
    for (int f = 0; f < nf; f++) {
        for (int k2 = 0; k2 < n; k2++)
            for (int a = 0; a < m; a++) {
                double s = 0;
                for (int k1 = 0; k1 < n; k1++) s += px[k1 + n * a] * c[(k1 + n * k2) * nf + f];
                w[a + m * k2] = s;
            }
        for (int b = 0; b < m; b++)
            for (int a = 0; a < m; a++) {
                double s = 0;
                for (int k2 = 0; k2 < n; k2++) s += py[k2 + n * b] * w[a + m * k2];
                out[(a + m * b) * nf + f] = s;
            }
    }
}

/* fun_err with iptype=2 (common/tree_vol_coeffs_nd.f:855-935): the largest
 * l2 norm over the nf functions of the masked coefficients, times rscale */
static double nt_fun_err(int nf, int npbox, const double *c, const double *rmask, double rscale) {
    double e = 0;
    for (int f = 0; f < nf; f++) {
        double s = 0;
        for (int i = 0; i < npbox; i++) s += rmask[i] * c[i * nf + f] * c[i * nf + f];
        s = sqrt(s);
        if (s > e) e = s;
    }
    return e * rscale;
}

static double nt_dist(double d, int iperiod, double bs0) {
    if (iperiod == 1) {
        double dp = d + bs0, dm = d - bs0;
        if (fabs(d) > fabs(dp)) d = dp;
        if (fabs(d) > fabs(dm)) d = dm;
    }
    return fabs(d);
}

/* box arrays moved together by the reorganisations */
typedef struct {
    tree_t *T;
    double *ctr, *pot, *coefs, *grad, *hess;
    int *iflag; /* may be NULL */
    int nf[4];  /* doubles per box of pot, coefs, grad, hess */
    int nbmax;
} nt_boxes;

/* Renumber: new position p takes old box order[p] (0-based); parents and
 * children are mapped through the inverse, -1 kept (vol_tree_reorg and its
 * variants, common/tree_vol_coeffs_nd.f:1973-2157, 2384-2570, 2916-3126). */
static void nt_permute(nt_boxes *B, const int *order, int cnt) {
    tree_t *T = B->T;
    int nbmax = B->nbmax;
    int *inv = (int *)malloc(sizeof(int) * nbmax);
    for (int i = 0; i < nbmax; i++) inv[i] = -1;
    for (int p = 0; p < cnt; p++) inv[order[p]] = p;
    int *il = (int *)malloc(sizeof(int) * nbmax), *ip = (int *)malloc(sizeof(int) * nbmax);
    int *nc = (int *)malloc(sizeof(int) * nbmax), *ic = (int *)malloc(sizeof(int) * 4 * nbmax);
    int *fl = (int *)malloc(sizeof(int) * nbmax);
    double *c2 = (double *)malloc(sizeof(double) * 2 * nbmax);
    memcpy(il, T->ilevel, sizeof(int) * nbmax);
    memcpy(ip, T->iparent, sizeof(int) * nbmax);
    memcpy(nc, T->nchild, sizeof(int) * nbmax);
    memcpy(ic, T->ichild, sizeof(int) * 4 * nbmax);
    memcpy(c2, B->ctr, sizeof(double) * 2 * nbmax);
    if (B->iflag) memcpy(fl, B->iflag, sizeof(int) * nbmax);
    for (int p = 0; p < cnt; p++) {
        int o = order[p];
        T->ilevel[p] = il[o];
        T->nchild[p] = nc[o];
        T->iparent[p] = ip[o] > 0 ? inv[ip[o] - 1] + 1 : -1;
        for (int j = 0; j < 4; j++) T->ichild[4 * p + j] = ic[4 * o + j] > 0 ? inv[ic[4 * o + j] - 1] + 1 : -1;
        B->ctr[2 * p] = c2[2 * o];
        B->ctr[2 * p + 1] = c2[2 * o + 1];
        if (B->iflag) B->iflag[p] = fl[o];
    }
    double *arr[4] = {B->pot, B->coefs, B->grad, B->hess};
    for (int a = 0; a < 4; a++) {
        if (!arr[a] || B->nf[a] == 0) continue;
        size_t w = (size_t)B->nf[a];
        double *cp = (double *)malloc(sizeof(double) * w * nbmax);
        memcpy(cp, arr[a], sizeof(double) * w * nbmax);
        for (int p = 0; p < cnt; p++) memcpy(arr[a] + w * p, cp + w * order[p], sizeof(double) * w);
        free(cp);
    }
    free(inv); free(il); free(ip); free(nc); free(ic); free(fl); free(c2);
}

/* vol_tree_refine_boxes_flag_interp (common/tree_vol_coeffs_nd.f:3127-3223):
 * the flagged boxes of [first, last] (0-based) are split in order; the
 * children's pot comes from the parent's coefficients, grad/hess from
 * theirs; children of flag-1 boxes become flag++ (3). */
static int nt_refine_flagged(nt_boxes *B, int first, int last, int ilev, int *nbctr, int nd, int n,
                             int ifpgh, const double *u, const double *p2c, const double *bsz) {
    tree_t *T = B->T;
    int npbox = n * n;
    double *gbuf = (double *)malloc(sizeof(double) * 5 * npbox * nd), *hbuf = gbuf + 2 * npbox * nd;
    for (int ib = first; ib <= last; ib++) {
        if (B->iflag[ib] <= 0) continue;
        if (*nbctr + 4 > B->nbmax) { free(gbuf); return 7; }
        double *gcoef = NULL, *hcoef = NULL;
        if (ifpgh >= 2) { nt_trans(2 * nd, n, u, B->grad + (size_t)ib * 2 * npbox * nd, gbuf); gcoef = gbuf; }
        if (ifpgh >= 3) { nt_trans(3 * nd, n, u, B->hess + (size_t)ib * 3 * npbox * nd, hbuf); hcoef = hbuf; }
        T->nchild[ib] = 4;
        double bsh = bsz[ilev + 1] / 2;
        for (int j = 0; j < 4; j++) {
            int jb = *nbctr + j;
            int sx = j & 1, sy = j >> 1;
            B->ctr[2 * jb] = B->ctr[2 * ib] + (sx ? bsh : -bsh);
            B->ctr[2 * jb + 1] = B->ctr[2 * ib + 1] + (sy ? bsh : -bsh);
            const double *px = p2c + sx * n * n, *py = p2c + sy * n * n;
            nt_eval(nd, n, B->coefs + (size_t)ib * npbox * nd, n, px, py, B->pot + (size_t)jb * npbox * nd);
            nt_trans(nd, n, u, B->pot + (size_t)jb * npbox * nd, B->coefs + (size_t)jb * npbox * nd);
            if (gcoef) nt_eval(2 * nd, n, gcoef, n, px, py, B->grad + (size_t)jb * 2 * npbox * nd);
            if (hcoef) nt_eval(3 * nd, n, hcoef, n, px, py, B->hess + (size_t)jb * 3 * npbox * nd);
            T->iparent[jb] = ib + 1;
            T->nchild[jb] = 0;
            for (int l = 0; l < 4; l++) T->ichild[4 * jb + l] = -1;
            T->ichild[4 * ib + j] = jb + 1;
            T->ilevel[jb] = ilev + 1;
            if (B->iflag[ib] == 1) B->iflag[jb] = 3;
            if (B->iflag[ib] == 2) B->iflag[jb] = 0;
        }
        *nbctr += 4;
    }
    free(gbuf);
    return 0;
}

/* vol_updateflags (common/tree_vol_coeffs_nd.f:1733-1850) on [first, last] */
static void nt_updateflags(nt_boxes *B, int first, int last, int curlev, int iperiod, const double *bsz) {
    tree_t *T = B->T;
    double distest = 1.05 * (bsz[curlev] + bsz[curlev + 1]) / 2.0, bs0 = bsz[0];
    for (int ib = first; ib <= last; ib++) {
        if (B->iflag[ib] != 3) continue;
        B->iflag[ib] = 0;
        for (int i = 0; i < T->ncoll[ib] && B->iflag[ib] == 0; i++) {
            int jb = T->coll[9 * ib + i] - 1;
            for (int j = 0; j < 4; j++) {
                int kb = T->ichild[4 * jb + j];
                if (kb <= 0) continue;
                kb -= 1;
                if (T->nchild[kb] <= 0) continue;
                int ict = 0;
                for (int k = 0; k < 2; k++)
                    if (nt_dist(B->ctr[2 * kb + k] - B->ctr[2 * ib + k], iperiod, bs0) <= distest) ict++;
                if (ict == 2) { B->iflag[ib] = 1; break; }
            }
        }
    }
}

/* vol_tree_reorg_pgh (common/tree_vol_coeffs_nd.f:2916-3126): levels 0 and
 * 1 stay, every deeper level lists its laddr range then its tail */
static void nt_reorg_tail(nt_boxes *B, int nlevels, const int *tail, int cnt) {
    tree_t *T = B->T;
    int *order = (int *)malloc(sizeof(int) * (cnt > 0 ? cnt : 1));
    int p = 0;
    for (int l = 0; l <= (nlevels < 1 ? nlevels : 1); l++)
        for (int b = LFIRST(T, l); b <= LLAST(T, l); b++) order[p++] = b;
    int *newl = (int *)malloc(sizeof(int) * 2 * (nlevels + 1));
    for (int l = 0; l <= nlevels; l++) { newl[2 * l] = T->laddr[2 * l]; newl[2 * l + 1] = T->laddr[2 * l + 1]; }
    for (int l = 2; l <= nlevels; l++) {
        newl[2 * l] = p + 1;
        for (int b = LFIRST(T, l); b <= LLAST(T, l); b++) order[p++] = b;
        for (int b = tail[2 * l] - 1; b <= tail[2 * l + 1] - 1; b++) order[p++] = b;
        newl[2 * l + 1] = p;
    }
    nt_permute(B, order, p);
    for (int l = 0; l <= nlevels; l++) { T->laddr[2 * l] = newl[2 * l]; T->laddr[2 * l + 1] = newl[2 * l + 1]; }
    free(order);
    free(newl);
}

static void nt_colleagues(tree_t *T, int nlevels, int iperiod) {
    T->nlevels = nlevels;
    T->nboxes = T->laddr[2 * nlevels + 1];
    oracle_computecoll(T, iperiod);
}

/* vol_tree_fix_lr_interp (common/tree_vol_coeffs_nd.f:2571-2915) */
static int nt_fix_lr(nt_boxes *B, int nlevels, int iperiod, int nd, int n, int ifpgh, const double *u,
                     const double *p2c, const double *bsz) {
    tree_t *T = B->T;
    double bs0 = bsz[0];
    int nboxes = T->laddr[2 * nlevels + 1];
    for (int i = 0; i < B->nbmax; i++) B->iflag[i] = 0;
    /* flag: leaves touching a box two levels finer (via the grandparent's colleagues) */
    for (int ilev = nlevels; ilev >= 2; ilev--) {
        double distest = 1.05 * (bsz[ilev - 1] + bsz[ilev - 2]) / 2.0;
        for (int ib = LFIRST(T, ilev); ib <= LLAST(T, ilev); ib++) {
            int dad = T->iparent[ib] - 1, gd = T->iparent[dad] - 1;
            for (int i = 0; i < T->ncoll[gd]; i++) {
                int jb = T->coll[9 * gd + i] - 1;
                if (T->nchild[jb] != 0 || B->iflag[jb] != 0) continue;
                int ict = 0;
                for (int k = 0; k < 2; k++)
                    if (nt_dist(B->ctr[2 * jb + k] - B->ctr[2 * dad + k], iperiod, bs0) <= distest) ict++;
                if (ict == 2) B->iflag[jb] = 1;
            }
        }
    }
    /* flag+: leaves bigger than and touching a flagged or flag+ box */
    for (int ilev = nlevels; ilev >= 1; ilev--) {
        double distest = 1.05 * (bsz[ilev] + bsz[ilev - 1]) / 2.0;
        for (int ib = LFIRST(T, ilev); ib <= LLAST(T, ilev); ib++) {
            if (B->iflag[ib] != 1 && B->iflag[ib] != 2) continue;
            int dad = T->iparent[ib] - 1;
            for (int i = 0; i < T->ncoll[dad]; i++) {
                int jb = T->coll[9 * dad + i] - 1;
                if (T->nchild[jb] != 0 || B->iflag[jb] != 0) continue;
                int ict = 0;
                for (int k = 0; k < 2; k++)
                    if (nt_dist(B->ctr[2 * jb + k] - B->ctr[2 * ib + k], iperiod, bs0) <= distest) ict++;
                if (ict == 2) B->iflag[jb] = 2;
            }
        }
    }
    int *tail = (int *)malloc(sizeof(int) * 2 * (nlevels + 1));
    for (int l = 0; l <= nlevels; l++) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
    int rc = 0;
    for (int ilev = 1; ilev <= nlevels - 2 && !rc; ilev++) {
        tail[2 * (ilev + 1)] = nboxes + 1;
        rc = nt_refine_flagged(B, LFIRST(T, ilev), LLAST(T, ilev), ilev, &nboxes, nd, n, ifpgh, u, p2c, bsz);
        tail[2 * (ilev + 1) + 1] = nboxes;
    }
    if (rc) { free(tail); return rc; }
    nt_reorg_tail(B, nlevels, tail, nboxes);
    nt_colleagues(T, nlevels, iperiod);
    for (int i = 0; i < nboxes; i++)
        if (B->iflag[i] != 3) B->iflag[i] = 0;
    for (int l = 0; l <= nlevels; l++) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
    for (int ilev = 2; ilev <= nlevels - 2 && !rc; ilev++) {
        nt_updateflags(B, LFIRST(T, ilev), LLAST(T, ilev), ilev, iperiod, bsz);
        nt_updateflags(B, tail[2 * ilev] - 1, tail[2 * ilev + 1] - 1, ilev, iperiod, bsz);
        tail[2 * (ilev + 1)] = nboxes + 1;
        rc = nt_refine_flagged(B, LFIRST(T, ilev), LLAST(T, ilev), ilev, &nboxes, nd, n, ifpgh, u, p2c, bsz);
        if (!rc)
            rc = nt_refine_flagged(B, tail[2 * ilev] - 1, tail[2 * ilev + 1] - 1, ilev, &nboxes, nd, n, ifpgh, u,
                                   p2c, bsz);
        tail[2 * (ilev + 1) + 1] = nboxes;
        /* colleagues of the boxes just created: children of the parent's
         * colleagues that touch */
        for (int ib = tail[2 * (ilev + 1)] - 1; ib <= tail[2 * (ilev + 1) + 1] - 1; ib++) {
            T->ncoll[ib] = 0;
            int dad = T->iparent[ib] - 1;
            for (int i = 0; i < T->ncoll[dad]; i++) {
                int jb = T->coll[9 * dad + i] - 1;
                for (int j = 0; j < 4; j++) {
                    int kb = T->ichild[4 * jb + j];
                    if (kb <= 0) continue;
                    kb -= 1;
                    int ok = 1;
                    for (int k = 0; k < 2; k++)
                        if (nt_dist(B->ctr[2 * kb + k] - B->ctr[2 * ib + k], iperiod, bs0) > 1.05 * bsz[ilev + 1]) {
                            ok = 0;
                            break;
                        }
                    if (ok) T->coll[9 * ib + T->ncoll[ib]++] = kb + 1;
                }
            }
        }
    }
    if (!rc) {
        nt_reorg_tail(B, nlevels, tail, nboxes);
        nt_colleagues(T, nlevels, iperiod);
    }
    free(tail);
    return rc;
}

static int newtree_steps(const newtree_in *S) {
    tree_t *T = S->T;
    const int nd = S->nd, n = S->n, npbox = n * n, nbmax = S->nbmax, ifpgh = S->ifpgh, iperiod = S->iperiod;
    const int mnl = S->mnl, nloc = loctab_half(S->mref), nslot = 2 * nloc + 1;
    const size_t nn = (size_t)n * n;
    int nlevels = *S->nlevels;
    const double *bsz = S->boxsize;
    const double bs0 = bsz[0];
    double *ctr = S->centers;
    double *pot = S->pot, *grad = S->grad, *hess = S->hess;
    const int H = (S->npw + 1) / 2, nexp = S->npw * H;
    int rc = 0;

    /* ===== STEP 8 (bfgt/boxfgt.f:1172-1412) ===== */
    double t1 = wtime();
    /* tens_prod_get_rmask, iptype=2 (common/polytens.f:1579-1621) */
    double rmask[MAXN * MAXN], rsum = 0;
    for (int k2 = 0; k2 < n; k2++)
        for (int k1 = 0; k1 < n; k1++) {
            int on = (k1 + 1) * (k1 + 1) + (k2 + 1) * (k2 + 1) >= n * n;
            rmask[k1 + n * k2] = on;
            rsum += on;
        }
    rsum = sqrt(rsum);
    /* treedata_lpnorm, iptype=2 (common/tree_data_routs_nd.f:1131-1225) */
    double rnorm = 0;
    for (int l = 0; l <= nlevels; l++) {
        double sc = (bsz[l] / 2) * (bsz[l] / 2);
        for (int b = LFIRST(T, l); b <= LLAST(T, l); b++) {
            if (T->nchild[b] != 0) continue;
            for (int i = 0; i < npbox; i++)
                for (int d = 0; d < nd; d++) {
                    double v = pot[((size_t)b * npbox + i) * nd + d];
                    rnorm += v * v * S->wq[i % n] * S->wq[i / n] * sc;
                }
        }
    }
    rnorm = sqrt(rnorm);
    const double reps = S->eps * sqrt(1.0 / (bs0 * bs0)) * rnorm;
    const int nboxes0 = T->nboxes;

    int *ifrefine = (int *)calloc(nbmax, sizeof(int));
    int *irefinelev = (int *)calloc(nbmax, sizeof(int));
    int *imaxrefinelev = (int *)calloc(nbmax, sizeof(int));
    int *iflag = (int *)calloc(nbmax, sizeof(int));
    int *pwexp = (int *)calloc(nbmax, sizeof(int));
    int *nl1 = (int *)calloc(nbmax, sizeof(int));
    int *l1 = (int *)calloc((size_t)nbmax * mnl, sizeof(int));
    double *coefsp = (double *)calloc((size_t)nbmax * npbox * nd, sizeof(double));
    cplx **locp = (cplx **)calloc(nbmax, sizeof(cplx *));
    cplx *locnew = S->exps ? (cplx *)calloc((size_t)(nbmax - nboxes0) * S->blk + 1, sizeof(cplx)) : NULL;
    int *nblock = (int *)calloc(nlevels + 2, sizeof(int));
    for (int b = 0; b < nboxes0; b++) {
        pwexp[b] = S->ifpwexp[b];
        nl1[b] = S->nlist1[b];
        memcpy(l1 + (size_t)mnl * b, S->list1 + (size_t)mnl * b, sizeof(int) * mnl);
        if (S->exps && S->nlevstart >= 0 && T->ilevel[b] >= S->nlevstart) locp[b] = S->exps + S->aloc[b];
    }
    /* coefficients of pot on every leaf (treedata_trans_nd) */
    for (int b = 0; b < nboxes0; b++)
        if (T->nchild[b] == 0) nt_trans(nd, n, S->umat, pot + (size_t)b * npbox * nd, coefsp + (size_t)b * npbox * nd);
    /* leaves above the finest level whose pot is not resolved */
    int maxrefinelev = 0;
    for (int l = 0; l <= nlevels - 1; l++) {
        double rscale = bsz[l] / 2;
        for (int b = LFIRST(T, l); b <= LLAST(T, l); b++) {
            if (T->nchild[b] != 0) continue;
            double erra = nt_fun_err(nd, npbox, coefsp + (size_t)b * npbox * nd, rmask, rscale) / rsum;
            if (erra > reps) {
                ifrefine[b] = 1;
                imaxrefinelev[b] = (int)ceil(log(erra / reps) / log(2.0) / n);
            }
            if (imaxrefinelev[b] > maxrefinelev) maxrefinelev = imaxrefinelev[b];
        }
    }
    /* refine; children are appended and refined in turn when the loop
     * reaches them.  Deviation: a box at the finest level is not split (the
     * reference would write a level its boxsize, laddr and tables do not
     * hold). */
    int nnew = 0, ncapped = 0;
    for (int ib = 0; ib < nbmax && !rc; ib++) {
        if (!(ifrefine[ib] == 1 && irefinelev[ib] <= maxrefinelev)) continue;
        int ilev = T->ilevel[ib];
        if (ilev >= nlevels) {
            ncapped++;
            continue;
        }
        irefinelev[ib]++;
        double bsh = bsz[ilev] / 4, rscale = bsh;
        T->nchild[ib] = 4;
        for (int j = 0; j < 4 && !rc; j++) {
            nnew++;
            int jb = nboxes0 + nnew - 1;
            if (jb >= nbmax) { rc = 7; break; }
            pwexp[jb] = pwexp[ib];
            nl1[jb] = nl1[ib];
            memcpy(l1 + (size_t)mnl * jb, l1 + (size_t)mnl * ib, sizeof(int) * mnl);
            irefinelev[jb] = irefinelev[ib] + 1;
            ctr[2 * jb] = ctr[2 * ib] + ((j & 1) ? bsh : -bsh);
            ctr[2 * jb + 1] = ctr[2 * ib + 1] + ((j >> 1) ? bsh : -bsh);
            T->iparent[jb] = ib + 1;
            T->nchild[jb] = 0;
            for (int l = 0; l < 4; l++) T->ichild[4 * jb + l] = -1;
            T->ichild[4 * ib + j] = jb + 1;
            int jlev = ilev + 1;
            T->ilevel[jb] = jlev;
            nblock[jlev]++;
            double *pj = pot + (size_t)jb * npbox * nd;
            double *gj = grad ? grad + (size_t)jb * npbox * nd * 2 : NULL;
            double *hj = hess ? hess + (size_t)jb * npbox * nd * 3 : NULL;
            if (pwexp[jb] == 1) {
                /* the parent's local expansion shifted to the child
                 * (gnd_shiftpw_loc), evaluated on its grid */
                int k = msshift_ind(ctr + 2 * jb, ctr + 2 * ib);
                int klev = nlevels - ilev;
                const cplx *w = S->wms + (size_t)(klev - 1) * 4 * nexp + (size_t)k * nexp;
                locp[jb] = locnew + (size_t)(jb - nboxes0) * S->blk;
                for (int d = 0; d < nd; d++)
                    for (int q = 0; q < nexp; q++)
                        locp[jb][(size_t)d * nexp + q] = locp[ib][(size_t)d * nexp + q] * w[q];
                pw2pgh(nd, n, S->npw, locp[jb], S->tp + (size_t)n * S->npw * jlev, S->tx + (size_t)n * S->npw * jlev,
                       S->txx + (size_t)n * S->npw * jlev, ifpgh, pj, gj, hj);
            }
            /* the parent's near-field sources (list 1 is symmetric) with the
             * child's level tables */
            for (int k = 0; k < nl1[jb]; k++) {
                int kb = l1[(size_t)mnl * jb + k];
                int ixy[2];
                loctab_ind(iperiod, ctr + 2 * jb, ctr + 2 * kb, bsz[T->ilevel[kb]], bs0, S->mref, ixy);
                if (abs(ixy[0]) > nloc || abs(ixy[1]) > nloc) { rc = 8; break; }
                tens_prod_to_pghloc(nd, n, S->fvals + (size_t)kb * npbox * nd, ifpgh, pj, gj, hj,
                                    S->tab + nn * nslot * jlev, S->tabx + nn * nslot * jlev,
                                    S->tabxx + nn * nslot * jlev, S->ind + 2 * (n + 1) * nslot * jlev, nloc, ixy);
            }
            nt_trans(nd, n, S->umat, pj, coefsp + (size_t)jb * npbox * nd);
            double erra = nt_fun_err(nd, npbox, coefsp + (size_t)jb * npbox * nd, rmask, rscale) / rsum;
            if (erra > reps) {
                ifrefine[jb] = 1;
                /* generations below the original leaf, against its estimate */
                int kb = jb, k, kdad = -1;
                for (k = 1; k <= maxrefinelev; k++) {
                    kdad = T->iparent[kb] - 1;
                    if (kdad < nboxes0) break;
                    kb = kdad;
                }
                if (kdad < 0 || k > imaxrefinelev[kdad]) ifrefine[jb] = 0;
            }
        }
    }
    if (S->info) {
        S->info[6] = nnew;
        S->info[7] = ncapped;
    }
    nt_boxes B = {T, ctr, pot, coefsp, ifpgh >= 2 ? grad : NULL, ifpgh >= 3 ? hess : NULL, iflag,
                  {npbox * nd, npbox * nd, ifpgh >= 2 ? 2 * npbox * nd : 0, ifpgh >= 3 ? 3 * npbox * nd : 0}, nbmax};
    if (!rc && nnew > 0) {
        /* fgt_vol_tree_reorg (common/tree_vol_coeffs_nd.f:1973-2157): per
         * level the old boxes, then the new ones in creation order */
        int cnt = nboxes0 + nnew;
        int *order = (int *)malloc(sizeof(int) * cnt);
        int p = 0;
        order[p++] = 0;
        for (int l = 1; l <= nlevels; l++) {
            int first = p + 1;
            for (int b = LFIRST(T, l); b <= LLAST(T, l); b++) order[p++] = b;
            for (int b = nboxes0; b < cnt; b++)
                if (T->ilevel[b] == l) order[p++] = b;
            T->laddr[2 * l] = first;
            T->laddr[2 * l + 1] = p;
        }
        nt_permute(&B, order, p);
        free(order);
        T->nboxes = p;
    }
    double t2 = wtime();
    S->timeinfo[8] += t2 - t1;

    if (!rc && nnew > 0 && !(nnew < 0.1 * nboxes0)) {
        /* ===== STEP 9: coarsen (vol_tree_coarsen, common/tree_vol_coeffs_nd.f:2158-2383) ===== */
        t1 = wtime();
        int n2 = n / 2;
        double c2p[2 * MAXN * MAXN], pv[MAXN + 2];
        for (int i = 0; i < n2; i++) {
            if (S->ipoly == 0) oracle_legepols(S->xq[i] * 2 + 1.0, n - 1, pv);
            else oracle_chebpols(S->xq[i] * 2 + 1.0, n - 1, pv);
            for (int k = 0; k < n; k++) c2p[k + n * i] = pv[k];
            if (S->ipoly == 0) oracle_legepols(S->xq[n2 + i] * 2 - 1.0, n - 1, pv);
            else oracle_chebpols(S->xq[n2 + i] * 2 - 1.0, n - 1, pv);
            for (int k = 0; k < n; k++) c2p[n * n + k + n * i] = pv[k];
        }
        int *ifdelete = (int *)calloc(nbmax, sizeof(int));
        int ndel = 0;
        double *v2 = (double *)malloc(sizeof(double) * 6 * npbox * nd), *cc = v2 + 3 * npbox * nd;
        for (int ilev = nlevels - 1; ilev >= 0; ilev--) {
            for (int ib = LFIRST(T, ilev); ib <= LLAST(T, ilev); ib++) {
                if (T->nchild[ib] != 4) continue;
                int nleaf = 0;
                for (int j = 0; j < 4; j++)
                    if (T->nchild[T->ichild[4 * ib + j] - 1] == 0) nleaf++;
                if (nleaf != 4) continue;
                double *pb = pot + (size_t)ib * npbox * nd;
                for (int j = 0; j < 4; j++) {
                    int jb = T->ichild[4 * ib + j] - 1;
                    const double *px = c2p + (j & 1) * n * n, *py = c2p + (j >> 1) * n * n;
                    nt_eval(nd, n, coefsp + (size_t)jb * npbox * nd, n2, px, py, v2);
                    int ox = (j & 1) * n2, oy = (j >> 1) * n2;
                    for (int b = 0; b < n2; b++)
                        for (int a = 0; a < n2; a++)
                            for (int d = 0; d < nd; d++)
                                pb[((ox + a) + n * (oy + b)) * nd + d] = v2[(a + n2 * b) * nd + d];
                }
                nt_trans(nd, n, S->umat, pb, coefsp + (size_t)ib * npbox * nd);
                double erra = nt_fun_err(nd, npbox, coefsp + (size_t)ib * npbox * nd, rmask, 1.0) / rsum;
                if (erra > reps) continue;
                for (int a = 0; a < 2; a++) {
                    int nf = (a == 0 ? 2 : 3) * nd;
                    double *arr = a == 0 ? (ifpgh >= 2 ? grad : NULL) : (ifpgh >= 3 ? hess : NULL);
                    if (!arr) continue;
                    for (int j = 0; j < 4; j++) {
                        int jb = T->ichild[4 * ib + j] - 1;
                        const double *px = c2p + (j & 1) * n * n, *py = c2p + (j >> 1) * n * n;
                        nt_trans(nf, n, S->umat, arr + (size_t)jb * npbox * nf, cc);
                        nt_eval(nf, n, cc, n2, px, py, v2);
                        int ox = (j & 1) * n2, oy = (j >> 1) * n2;
                        for (int b = 0; b < n2; b++)
                            for (int a2 = 0; a2 < n2; a2++)
                                for (int f = 0; f < nf; f++)
                                    arr[((size_t)ib * npbox + (ox + a2) + n * (oy + b)) * nf + f] =
                                        v2[(a2 + n2 * b) * nf + f];
                    }
                }
                for (int j = 0; j < 4; j++) {
                    int jb = T->ichild[4 * ib + j] - 1;
                    ndel++;
                    T->iparent[jb] = -1;
                    T->nchild[jb] = 0;
                    for (int l = 0; l < 4; l++) T->ichild[4 * jb + l] = -1;
                    ifdelete[jb] = 1;
                }
                T->nchild[ib] = 0;
                for (int j = 0; j < 4; j++) T->ichild[4 * ib + j] = -1;
            }
        }
        if (ndel > 0) {
            /* fgt_vol_tree_reorg_after_coarsen (common/tree_vol_coeffs_nd.f:2384-2570) */
            int cnt0 = T->laddr[2 * nlevels + 1];
            int *order = (int *)malloc(sizeof(int) * cnt0);
            int p = 0;
            order[p++] = 0;
            for (int l = 1; l <= nlevels; l++) {
                int first = p + 1;
                for (int b = LFIRST(T, l); b <= LLAST(T, l); b++)
                    if (!ifdelete[b]) order[p++] = b;
                T->laddr[2 * l] = first;
                T->laddr[2 * l + 1] = p;
            }
            nt_permute(&B, order, p);
            free(order);
            int nnl = nlevels;
            for (int l = nlevels; l >= 0; l--)
                if (T->laddr[2 * l + 1] < T->laddr[2 * l]) nnl--;
            nlevels = nnl;
            T->nlevels = nlevels;
            T->nboxes = T->laddr[2 * nlevels + 1];
        }
        free(ifdelete);
        free(v2);
        t2 = wtime();
        S->timeinfo[9] += t2 - t1;
    }

    if (!rc && nnew > 0) {
        /* ===== STEP 10: 2:1 rebalance (bfgt/boxfgt.f:1459-1483) ===== */
        t1 = wtime();
        nt_colleagues(T, nlevels, iperiod);
        double p2c[2 * MAXN * MAXN], pv[MAXN + 2];
        for (int i = 0; i < n; i++) {
            if (S->ipoly == 0) oracle_legepols(S->xq[i] / 2 - 0.5, n - 1, pv);
            else oracle_chebpols(S->xq[i] / 2 - 0.5, n - 1, pv);
            for (int k = 0; k < n; k++) p2c[k + n * i] = pv[k];
            if (S->ipoly == 0) oracle_legepols(S->xq[i] / 2 + 0.5, n - 1, pv);
            else oracle_chebpols(S->xq[i] / 2 + 0.5, n - 1, pv);
            for (int k = 0; k < n; k++) p2c[n * n + k + n * i] = pv[k];
        }
        rc = nt_fix_lr(&B, nlevels, iperiod, nd, n, ifpgh, S->umat, p2c, bsz);
        t2 = wtime();
        S->timeinfo[10] += t2 - t1;
    }
    T->nlevels = nlevels;
    T->nboxes = T->laddr[2 * nlevels + 1];
    *S->nlevels = nlevels;
    free(ifrefine); free(irefinelev); free(imaxrefinelev); free(iflag); free(pwexp); free(nl1); free(l1);
    free(coefsp); free(locp); free(locnew); free(nblock);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Extra targets: bfgtmain STEP 11 with iffast=0 (bfgt/boxfgt.f:1494-1563) */
/* ------------------------------------------------------------------ */

/* Leaf of the point (x, y): descend from the root choosing the child on the
 * plus side of the centre when the offset is >= 0 (sort_pts_to_children and
 * find_childbox_ind, common/pts_tree_nd.f:650-741, driven level by level by
 * pts_tree_sort :1226-1253).  Child order mm, pm, mp, pp. */
static int locate_leaf(const tree_t *T, double x, double y) {
    int b = 0;
    while (T->nchild[b] > 0) {
        int k = (x - T->centers[2 * b] >= 0 ? 1 : 0) + (y - T->centers[2 * b + 1] >= 0 ? 2 : 0);
        b = T->ichild[4 * b + k] - 1;
    }
    return b;
}

/* One channel set (nf functions per point) of one leaf: values -> tensor
 * coefficients (ortho_trans_2d itype=0, common/tensor_prod_routs_nd.f:103-180:
 * x transform, then y transform), then the expansion at the scaled target
 * (ortho_evalt_2d: sum over x degree first, then y). */
static void eval_leaf(int nf, int n, int ipoly, const double *umat, const double *vals, double xs, double ys,
                      double *out) {
    double px[MAXN], py[MAXN];
    if (ipoly == 0) {
        oracle_legepols(xs, n - 1, px);
        oracle_legepols(ys, n - 1, py);
    } else {
        oracle_chebpols(xs, n - 1, px);
        oracle_chebpols(ys, n - 1, py);
    }
    double *work = (double *)malloc(sizeof(double) * nf * n * n);
    double *coef = (double *)malloc(sizeof(double) * nf * n * n);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++)
            for (int f = 0; f < nf; f++) {
                double dd = 0;
                for (int k1 = 0; k1 < n; k1++) dd += umat[k + n * k1] * vals[f + nf * (k1 + n * j)];
                work[f + nf * (k + n * j)] = dd;
            }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++)
            for (int f = 0; f < nf; f++) {
                double dd = 0;
                for (int i1 = 0; i1 < n; i1++) dd += umat[i + n * i1] * work[f + nf * (k + n * i1)];
                coef[f + nf * (k + n * i)] = dd;
            }
    for (int f = 0; f < nf; f++) {
        double pot = 0;
        for (int j = 0; j < n; j++) {
            double pp = 0;
            for (int k = 0; k < n; k++) pp += coef[f + nf * (k + n * j)] * px[k];
            pot += pp * py[j];
        }
        out[f] = pot;
    }
    free(work);
    free(coef);
}

/*
 * oracle_eval_targets: the leaf arrays pot (nd,npbox,*), grad (nd,2,npbox,*),
 * hess (nd,3,npbox,*) interpolated at ntarg targets (2,ntarg) by their
 * leaves' tensor polynomial expansions (treedata_trans_nd + treedata_evalt_nd,
 * common/tree_data_routs_nd.f:16-90, 710-806); pote (nd,ntarg),
 * grade (nd,2,ntarg), hesse (nd,3,ntarg) are overwritten for the channels
 * ifpghtarg asks for.
 */
int oracle_eval_targets(int nd, int ipoly, int norder, int nlevels, int *itree, const int *iptr,
                        const double *centers, const double *boxsize, int ntarg, const double *targs,
                        int ifpghtarg, const double *pot, const double *grad, const double *hess, double *pote,
                        double *grade, double *hesse) {
    if (norder < 1 || norder > MAXN) return 3;
    if (ifpghtarg < 1 || ifpghtarg > 3) return 5;
    int n = norder, npbox = n * n;
    tree_t T;
    tree_view(&T, nlevels, itree, iptr, centers, boxsize);
    double xq[MAXN], wq[MAXN], umat[MAXN * MAXN];
    if (ipoly == 0) oracle_legeexps(n, xq, wq, umat);
    else oracle_chebexps(n, xq, wq, umat);
    for (int it = 0; it < ntarg; it++) {
        double x = targs[2 * it], y = targs[2 * it + 1];
        int b = locate_leaf(&T, x, y);
        double sc = 2.0 / boxsize[T.ilevel[b]];
        double xs = (x - centers[2 * b]) * sc, ys = (y - centers[2 * b + 1]) * sc;
        eval_leaf(nd, n, ipoly, umat, pot + (size_t)b * npbox * nd, xs, ys, pote + (size_t)it * nd);
        if (ifpghtarg >= 2)
            eval_leaf(2 * nd, n, ipoly, umat, grad + (size_t)b * npbox * 2 * nd, xs, ys, grade + (size_t)it * 2 * nd);
        if (ifpghtarg >= 3)
            eval_leaf(3 * nd, n, ipoly, umat, hess + (size_t)b * npbox * 3 * nd, xs, ys, hesse + (size_t)it * 3 * nd);
    }
    return 0;
}
