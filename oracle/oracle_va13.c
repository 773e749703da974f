/*
 * oracle_va13.c — CPU restatement of the reference's VA13 variable-metric
 * minimizer (src/apps/em/minfor.f:1-182 with the factorized-Hessian routines
 * of src/apps/em/mc11a.f:1-143), for any number of variables.  TEST
 * INFRASTRUCTURE ONLY (the checker of the device minimizer).
 *
 * The Hessian approximation is kept as L D L^T in packed storage: for
 * i = 0..n-1, D_i followed by the n-1-i entries L_{i+1..n-1, i}.  All
 * arithmetic in double, in the order of the Fortran source; 0-based indices.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dmin(double a, double b) { return b < a ? b : a; }
static double dmax(double a, double b) { return b > a ? b : a; }

/* mc11a.f:1-124 — update of the factors of H by sig * z z^T (mk = 1: w holds
 * L^-1 z from the previous solve; mk = 0: computed here).  z and w may be the
 * same array, as in minfor's second update. */
static void va13_update(double* a, int n, double* z, double sig, double* w, int* ir, int mk,
                        double eps) {
    if (n <= 1) {
        a[0] = a[0] + sig * z[0] * z[0];
        *ir = 1;
        if (a[0] > 0.0) return;
        a[0] = 0.0;
        *ir = 0;
        return;
    }
    const int np = n + 1;
    double ti = 0.0, tim = 0.0;
    int mm;
    int ij = 0; /* 0-based packed position */
    if (sig > 0.0) goto rank_one_plus;
    if (sig == 0.0 || *ir == 0) return;
    ti = 1.0 / sig;
    if (mk != 0) {
        for (int i = 0; i < n; ++i) {
            if (a[ij] != 0.0) ti = ti + w[i] * w[i] / a[ij];
            ij += np - (i + 1);
        }
    } else {
        for (int i = 0; i < n; ++i) w[i] = z[i];
        for (int i = 0; i < n; ++i) {
            const double v = w[i];
            if (!(a[ij] > 0.0)) {
                w[i] = 0.0;
                ij += np - (i + 1);
                continue;
            }
            ti = ti + v * v / a[ij];
            for (int j = i + 1; j < n; ++j) {
                ++ij;
                w[j] = w[j] - v * a[ij];
            }
            ++ij;
        }
    }
    /* mc11a.f label 20 */
    if (*ir <= 0) {
        ti = 0.0;
        *ir = -*ir - 1;
    } else if (ti > 0.0) {
        ti = eps / sig;
        if (eps == 0.0) *ir = *ir - 1;
    } else if (mk - 1 <= 0) {
        goto rank_one_plus;
    }
    /* label 23: backward recurrence of the partial sums */
    mm = 1;
    tim = ti;
    for (int i = 1; i <= n; ++i) {
        const int j = np - i - 1; /* 0-based index of Fortran j = np - i */
        ij -= i;
        if (a[ij] != 0.0) tim = ti - w[j] * w[j] / a[ij];
        w[j] = ti;
        ti = tim;
    }
    goto apply;
rank_one_plus:
    mm = 0;
    tim = 1.0 / sig;
apply:
    /* label 41 */
    ij = 0;
    for (int i = 0; i < n; ++i) {
        const double v = z[i];
        if (!(a[ij] > 0.0)) {
            if (*ir > 0 || sig < 0.0 || v == 0.0) {
                ti = tim;
                ij += np - (i + 1);
                continue;
            }
            *ir = 1 - *ir;
            a[ij] = v * v / tim;
            if (i == n - 1) return;
            for (int j = i + 1; j < n; ++j) {
                ++ij;
                a[ij] = z[j] / v;
            }
            return;
        }
        const double al = v / a[ij];
        if (mm <= 0) ti = tim + v * al;
        else ti = w[i];
        const double r = ti / tim;
        a[ij] = a[ij] * r;
        if (r == 0.0 || i == n - 1) break;
        const double b = al / ti;
        if (r > 4.0) {
            const double gm = tim / ti;
            for (int j = i + 1; j < n; ++j) {
                ++ij;
                const double y = a[ij];
                a[ij] = b * z[j] + y * gm;
                z[j] = z[j] - v * y;
            }
        } else {
            for (int j = i + 1; j < n; ++j) {
                ++ij;
                z[j] = z[j] - v * a[ij];
                a[ij] = a[ij] + b * z[j];
            }
        }
        tim = ti;
        ++ij;
    }
    if (*ir < 0) *ir = -*ir;
}

/* mc11a.f:126-143 (entry mc11e) — z <- H^-1 z with H = L D L^T; w receives
 * L^-1 z.  Nothing is done while the factorization is rank deficient. */
static void va13_solve(const double* a, int n, double* z, double* w, int ir) {
    if (ir < n) return;
    w[0] = z[0];
    if (n <= 1) {
        z[0] = z[0] / a[0];
        return;
    }
    int ij = 0;
    for (int i = 1; i < n; ++i) {
        ij = i;
        double v = z[i];
        for (int j = 0; j < i; ++j) {
            v = v - a[ij] * z[j];
            ij += n - (j + 1);
        }
        w[i] = v;
        z[i] = v;
    }
    /* ij is the position of D_{n-1} */
    z[n - 1] = z[n - 1] / a[ij];
    for (int nip = 2; nip <= n; ++nip) {
        const int i = n - nip; /* 0-based */
        int ii = ij - nip;
        double v = z[i] / a[ii];
        ij = ii;
        for (int j = i + 1; j < n; ++j) {
            ++ii;
            v = v - a[ii] * z[j];
        }
        z[i] = v;
    }
}

/* The factor update alone, for tests of its branches. */
void orc_va13_update(double* a, int n, double* z, double sig, double* w, int* ir, int mk, double eps) {
    va13_update(a, n, z, sig, w, ir, mk, eps);
}

/* minfor.f:1-182.  fg(ctx, x, &f, g) evaluates the objective; the result is
 * the last evaluated point and value (what the reference's solver reports,
 * VA13Solver.cpp:48-69) and the lowest value seen. */
int orc_va13(int n, const double* x0, int ivmax, orc_fg fg, void* ctx, double* xlast,
             double* flast, double* fbest, int* nfun_out) {
    double *buf = (double*)calloc((size_t)(9 * n + n * (n + 1) / 2), sizeof(double));
    double *x = buf, *xaa = x + n, *xbb = xaa + n, *g = xbb + n, *ga = g + n, *gb = ga + n,
           *d = gb + n, *w = d + n, *ev = w + n, *h = ev + n;
    const double acc = 1e-9;
    double gesa = 0, fa = 0, fb = 0, dff = 0, dga = 0, dgb = 0, c = 1.0, stmin = 0, stepbd = 0,
           steplb = 0, fmn = 0, gmin = 0, step = 0;
    int isfv = 0, itr = 0, nfun = 0, ir = 0;
    for (int i = 0; i < n; ++i) x[i] = xaa[i] = x0[i];
    /* diagonal initial Hessian 0.01 */
    for (int i = 0, k = 0; i < n; ++i) {
        h[k] = 0.01 * c;
        k += n - i;
    }
    ir = n; /* rank of the factorization (minfor passes jn to mc11e) */
    fg(ctx, xaa, &gesa, g);
    memcpy(ev, xaa, sizeof(double) * n);
    double elast = gesa;
    for (;;) {
        /* label 110 */
        fa = gesa;
        isfv = 1;
        for (int i = 0; i < n; ++i) ga[i] = g[i];
        int restart = 0;
        while (!restart) {
            /* label 135: search direction */
            ++itr;
            for (int i = 0; i < n; ++i) d[i] = -ga[i];
            va13_solve(h, n, d, w, n);
            c = 0.0;
            dga = 0.0;
            for (int i = 0; i < n; ++i) {
                c = dmax(c, fabs(d[i]));
                dga = dga + ga[i] * d[i];
            }
            if (dga >= 0.0) goto label240;
            stmin = 0.0;
            stepbd = 0.0;
            steplb = acc / c;
            fmn = fa;
            gmin = dga;
            step = 1.0;
            if (dff <= 0.0) step = dmin(step, 1.0 / c);
            if (dff > 0.0) step = dmin(step, (dff + dff) / (-dga));
            for (;;) {
                /* label 170: evaluate at xaa + c d */
                c = stmin + step;
                if (nfun >= ivmax) goto done;
                ++nfun;
                for (int i = 0; i < n; ++i) xbb[i] = xaa[i] + c * d[i];
                fg(ctx, xbb, &fb, gb);
                memcpy(ev, xbb, sizeof(double) * n);
                elast = fb;
                if (isfv > 2) isfv = 2;
                int better = 0;
                if (fb < gesa) {
                    better = 1;
                } else if (!(fb > gesa)) {
                    double gl1 = 0.0, gl2 = 0.0;
                    for (int i = 0; i < n; ++i) {
                        gl1 = gl1 + g[i] * g[i];
                        gl2 = gl2 + gb[i] * gb[i];
                    }
                    if (gl2 < gl1) better = 1;
                }
                if (better) {
                    /* label 200 */
                    isfv = 3;
                    gesa = fb;
                    for (int i = 0; i < n; ++i) {
                        x[i] = xbb[i];
                        g[i] = gb[i];
                    }
                }
                /* label 220 */
                dgb = 0.0;
                for (int i = 0; i < n; ++i) dgb = dgb + gb[i] * d[i];
                if (fb - fa <= 0.1 * c * dga) {
                    /* label 280: new lower bound on the step */
                    stepbd = stepbd - step;
                    stmin = c;
                    fmn = fb;
                    gmin = dgb;
                    step = 9.0 * stmin;
                    if (stepbd > 0.0) step = 0.5 * stepbd;
                    c = dga + 3.0 * dgb - 4.0 * (fb - fa) / stmin;
                    if (c > 0.0) step = dmin(step, stmin * dmax(1.0, -dgb / c));
                    if (dgb < 0.7 * dga) continue; /* to 170 */
                    isfv = 4 - isfv;
                    if (stmin + step <= steplb) goto label240;
                    /* revise the factors: H + ga ga^T / dga + y y^T / (stmin (dgb - dga)) */
                    ir = -n;
                    for (int i = 0; i < n; ++i) {
                        xaa[i] = xbb[i];
                        xbb[i] = ga[i];
                        d[i] = gb[i] - ga[i];
                        ga[i] = gb[i];
                    }
                    va13_update(h, n, xbb, 1.0 / dga, w, &ir, 1, 0.0);
                    ir = -ir;
                    va13_update(h, n, d, 1.0 / (stmin * (dgb - dga)), d, &ir, 0, 0.0);
                    if (ir < n) goto done;
                    dff = fa - fb;
                    fa = fb;
                    break; /* to 135 */
                }
                if (step > steplb) {
                    /* label 270: cubic interpolation */
                    stepbd = step;
                    c = gmin + dgb - 3.0 * (fb - fmn) / step;
                    c = gmin / (c + gmin - sqrt(c * c - gmin * dgb));
                    step = step * dmax(0.1, c);
                    continue; /* to 170 */
                }
                goto label240;
            }
            continue;
        label240:
            if (isfv >= 2) {
                restart = 1; /* to 110 */
                continue;
            }
            goto done;
        }
    }
done:
    if (xlast) memcpy(xlast, ev, sizeof(double) * n);
    if (flast) *flast = elast;
    if (fbest) *fbest = gesa;
    if (nfun_out) *nfun_out = nfun;
    free(buf);
    return 0;
}
