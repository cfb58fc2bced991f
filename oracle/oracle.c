/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the reference
 * tracking algorithm, used by tests/, the smoke check and bench.py's
 * cpu_baseline leg as the checker. The product path never calls it.
 *
 * Each function restates one reference routine with the same floating-point
 * operation order, so that, built with -O2 -ffp-contract=off for baseline
 * x86-64 (as the reference is: setup.py:48-53, -std=c99), it reproduces the
 * reference bit for bit. Particles are independent on this path, so the
 * driver runs particle-outer (one particle through all turns and elements)
 * instead of at.c's element-outer loop; the per-element loss check and the
 * refpt copies happen at the same points of the sequence, which makes the
 * two orders equivalent.
 *
 * Parity pinning (tests/test_oracle.py): every pass method against the
 * reference integrators compiled from /root/reference (oracle/_ref) on
 * random and edge-case inputs, via the committed golden fixtures; the
 * reference's known-answer tests (pyat/test/test_basic_elements.py,
 * test_atpass.py); and closed forms (drift, thin-lens limits, symplecticity).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Forest-Ruth 4th-order coefficients: atintegrators/atconstants.h:15-19 */
#define OR_CBRT2 1.25992104989487316477
static const double OR_D1 = 1.0 / (2.0 * (2.0 - OR_CBRT2));
static const double OR_D2 = 0.5 - (1.0 / (2.0 * (2.0 - OR_CBRT2)));
static const double OR_K1 = 1.0 / (2.0 - OR_CBRT2);
static const double OR_K2 = 1.0 - 2.0 / (2.0 - OR_CBRT2);
/* atconstants.h:23-31 */
static const double OR_C0 = 2.99792458e8;
static const double OR_TWOPI = 2.0 * 3.14159265358979323846;
#define OR_RE 2.8179403205e-15
#define OR_E0_GEV 510.99895069e-6
static const double OR_RAD_CONST = 2.0 / 3.0 * OR_RE;

static char or_errbuf[256];

const char *oracle_last_error(void) { return or_errbuf; }

static int or_fail(int code, const char *msg)
{
    snprintf(or_errbuf, sizeof(or_errbuf), "%s", msg);
    return code;
}

/* ---------------------------------------------------------------------- */
/* Vector helpers: atlalib.c:17-75                                         */

static void or_add_t(double *r, const double *t)
{
    for (int i = 0; i < 6; i++) r[i] += t[i];
}

/* r <- M r, M column-major M[i+6j] (atlalib.c:17-32) */
static void or_mul_r(double *r, const double *m)
{
    double acc[6];
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
        for (int j = 0; j < 6; j++) s += m[i + 6 * j] * r[j];
        acc[i] = s;
    }
    memcpy(r, acc, sizeof(acc));
}

/* atlalib.c:70-75 */
static void or_change_pref(double *r, double s)
{
    r[1] /= s;
    r[3] /= s;
    r[4] = (r[4] + 1.0 - s) / s;
}

/* atlalib.c:59-68 and driftkickrad.c:36-45 (identical arithmetic) */
static void or_drift6(double *r, double len)
{
    double pn = 1 / (1 + r[4]);
    double nl = len * pn;
    r[0] += nl * r[1];
    r[2] += nl * r[3];
    r[5] += nl * pn * (r[1] * r[1] + r[3] * r[3]) / 2;
}

/* atlalib.c:169-175: strict inequalities, NaN compares false */
static void or_rect_ap(double *r, const double *lim)
{
    if (r[0] < lim[0] || r[0] > lim[1])
        r[5] = INFINITY;
    else if (r[2] < lim[2] || r[2] > lim[3])
        r[5] = INFINITY;
}

/* atlalib.c:177-182 */
static void or_ellip_ap(double *r, const double *ax)
{
    double u = r[0] / ax[0];
    double v = r[2] / ax[1];
    if ((u * u + v * v) >= 1) r[5] = INFINITY;
}

static void or_apertures(double *r, const oracle_elem *e)
{
    if (e->RApertures) or_rect_ap(r, e->RApertures);
    if (e->EApertures) or_ellip_ap(r, e->EApertures);
}

static void or_entry_transform(double *r, const oracle_elem *e)
{
    if (e->T1) or_add_t(r, e->T1);
    if (e->R1) or_mul_r(r, e->R1);
}

static void or_exit_transform(double *r, const oracle_elem *e)
{
    if (e->R2) or_mul_r(r, e->R2);
    if (e->T2) or_add_t(r, e->T2);
}

/* ---------------------------------------------------------------------- */
/* Drift and thin kicks: driftkick.c:37-115                                */

/* driftkick.c:37-48: nl = L/(1+delta) hoisted by the caller */
static void or_fastdrift(double *r, double nl)
{
    r[0] += nl * r[1];
    r[2] += nl * r[3];
    r[5] += nl * (r[1] * r[1] + r[3] * r[3]) / (2 * (1 + r[4]));
}

/* Horner evaluation of (B + iA)(x + iy)^n, driftkick.c:92-110 */
static void or_field(const double *r, const double *a, const double *b, int n,
                     double *re_out, double *im_out)
{
    double re = b[n];
    double im = a[n];
    for (int k = n - 1; k >= 0; k--) {
        double t = re * r[0] - im * r[2] + b[k];
        im = im * r[0] + re * r[2] + a[k];
        re = t;
    }
    *re_out = re;
    *im_out = im;
}

/* driftkick.c:92-115 */
static void or_str_kick(double *r, const double *a, const double *b, double kl, int n)
{
    double re, im;
    or_field(r, a, b, n, &re, &im);
    r[1] -= kl * re;
    r[3] += kl * im;
}

/* driftkick.c:51-89 */
static void or_bnd_kick(double *r, const double *a, const double *b, double kl,
                        double irho, int n)
{
    double re, im;
    or_field(r, a, b, n, &re, &im);
    r[1] -= kl * (re - (r[4] - r[0] * irho) * irho);
    r[3] += kl * im;
    r[5] += kl * irho * r[0];
}

/* ---------------------------------------------------------------------- */
/* Radiating drift and kicks: diff_drift.c:30-43, diff_str_kick.c:19-83,  */
/* diff_bnd_kick.c:85-174 (tracking only: bdiff == NULL)                   */

static void or_rad_drift(double *r, double len)
{
    double pn = 1.0 / (1.0 + r[4]);
    double nl = len * pn;
    r[0] += nl * r[1];
    r[2] += nl * r[3];
    r[5] += nl * pn * (r[1] * r[1] + r[3] * r[3]) / 2.0;
}

static void or_rad_str_kick(double *r, const double *a, const double *b, int n,
                            double kl, double rad_const)
{
    double re, im;
    or_field(r, a, b, n, &re, &im);
    double pn = 1.0 / (1.0 + r[4]);
    double xpr = r[1] * pn;
    double ypr = r[3] * pn;
    /* |e x B|^2 with bx = im, by = re (diff_str_kick.c:1-17) */
    double vn2 = 1.0 / (1.0 + xpr * xpr + ypr * ypr);
    double cross = im * ypr - re * xpr;
    double b2p = (re * re + im * im + cross * cross) * vn2;
    double factor = (1.0 + (xpr * xpr + ypr * ypr) / 2.0) / (pn * pn) * kl;
    r[4] -= rad_const * b2p * factor;
    pn = 1.0 / (1.0 + r[4]);
    r[1] = xpr / pn;
    r[3] = ypr / pn;
    r[1] -= kl * re;
    r[3] += kl * im;
}

static void or_rad_bnd_kick(double *r, const double *a, const double *b, int n,
                            double kl, double irho, double rad_const)
{
    double re, im;
    or_field(r, a, b, n, &re, &im);
    double pn = 1.0 / (1.0 + r[4]);
    double x = r[0];
    double xpr = r[1] * pn;
    double ypr = r[3] * pn;
    /* curved-frame |e x B|^2 with the full field by = re + irho */
    double by = re + irho;
    double bx = im;
    double hx = 1.0 + x * irho;
    double vn2 = 1.0 / (hx * hx + xpr * xpr + ypr * ypr);
    double h1 = 1 + x * irho;
    double byh = by * h1;
    double bxh = bx * h1;
    double cross = bx * ypr - by * xpr;
    double b2p = (byh * byh + bxh * bxh + cross * cross) * vn2;
    double factor = (1.0 + x * irho + (xpr * xpr + ypr * ypr) / 2.0) / (pn * pn) * kl;
    double d0 = r[4];
    r[4] -= rad_const * b2p * factor;
    pn = 1.0 / (1.0 + r[4]);
    r[1] = xpr / pn;
    r[3] = ypr / pn;
    r[1] -= kl * (re - (d0 - r[0] * irho) * irho);
    r[3] += kl * im;
    r[5] += kl * irho * r[0];
}

/* ---------------------------------------------------------------------- */
/* Edge focusing: atphyslib.c:15-79 (non-radiating) and                   */
/* diff_bend_fringe.c:2-53 (radiating, B propagation omitted)              */

/* exactdrift.c:4-16 (ExactDriftPass, ExactMultipolePass): absolute path */
static void or_exact_drift(double *r, double L)
{
    double pz = sqrt((1 + r[4]) * (1 + r[4]) - r[1] * r[1] - r[3] * r[3]);
    double nl = L / pz;
    r[0] += r[1] * nl;
    r[2] += r[3] * nl;
    r[5] += nl * (1.0 + r[4]);
}

/* diff_exactdrift.c:30-43 (ExactMultipoleRadPass), bdiff == NULL */
static void or_exact_rad_drift(double *r, double L)
{
    double pn = 1.0 / (1.0 + r[4]);
    double xpr = r[1] * pn;
    double ypr = r[3] * pn;
    double nl = L * pn / sqrt(1.0 - xpr * xpr - ypr * ypr);
    r[0] += nl * r[1];
    r[2] += nl * r[3];
    r[5] += nl * (1.0 + r[4]);
}

/* diff_str_exactkick.c:14-71, bdiff == NULL */
static void or_exact_rad_kick(double *r, const double *a, const double *b, int n, double L,
                              double rad_const)
{
    double re, im;
    or_field(r, a, b, n, &re, &im);
    double pn = 1.0 / (1.0 + r[4]);
    double xpr = r[1] * pn;
    double ypr = r[3] * pn;
    /* B2perp(bx = ImSum, by = ReSum): bx^2 + by^2 - (bx xpr + by ypr)^2 */
    double t = im * xpr + re * ypr;
    double b2p = im * im + re * re - t * t;
    double factor = L / (pn * pn) / sqrt(1.0 - xpr * xpr - ypr * ypr);
    r[4] -= rad_const * b2p * factor;
    pn = 1.0 / (1.0 + r[4]);
    r[1] = xpr / pn;
    r[3] = ypr / pn;
    r[1] -= L * re;
    r[3] += L * im;
}

/* exactmultipolefringe.c:2-95 (PTC multipole fringe, Forest 13.29),
 * skip_b0 = 0; edge = +1 at the entrance, -1 at the exit */
static void or_multipole_fringe(double *r, const double *pa, const double *pb, int max_order,
                                double edge)
{
    double fx = 0, fy = 0, fx_x = 0, fx_y = 0, fy_x = 0, fy_y = 0;
    double rx = 1.0, ix = 0.0;
    for (int n = 0; n <= max_order; n++) {
        double bn = pb[n], an = pa[n];
        int j = n + 1;
        double drx = rx, dix = ix;
        rx = drx * r[0] - dix * r[2];
        ix = drx * r[2] + dix * r[0];
        double u = bn * rx - an * ix;
        double v = bn * ix + an * rx;
        double du = bn * drx - an * dix;
        double dv = bn * dix + an * drx;
        double f1 = -edge / 4.0 / (j + 1);
        u = u * f1;
        v = v * f1;
        du = du * f1;
        dv = dv * f1;
        double dux = j * du, dvx = j * dv, duy = -j * dv, dvy = j * du;
        double nf = 1.0 * (j + 2) / j;
        fx += u * r[0] + nf * v * r[2];
        fy += u * r[2] - nf * v * r[0];
        fx_x += dux * r[0] + u + nf * r[2] * dvx;
        fx_y += duy * r[0] + nf * v + nf * r[2] * dvy;
        fy_x += dux * r[2] - nf * v - nf * r[0] * dvx;
        fy_y += duy * r[2] + u - nf * r[0] * dvy;
    }
    double del = 1.0 / (1 + r[4]);
    double ma = 1 - fx_x * del;
    double mb = -fy_x * del;
    double md = 1 - fy_y * del;
    double mc = -fx_y * del;
    r[0] = r[0] - fx * del;
    r[2] = r[2] - fy * del;
    double pxf = (md * r[1] - mb * r[3]) / (ma * md - mb * mc);
    double pyf = (ma * r[3] - mc * r[1]) / (ma * md - mb * mc);
    r[3] = pyf;
    r[1] = pxf;
    r[5] = r[5] - (r[1] * fx + r[3] * fy) * del * del;
}

static double or_fringe_corr(double irho, double edge, double fint, double gap, int method)
{
    if ((fint == 0.0) || (gap == 0.0) || (method == 0)) return 0.0;
    double s = sin(edge);
    double c = cos(edge);
    return irho * gap * fint * (1 + s * s) / c;
}

static void or_edge(double *r, double irho, double edge, double fint, double gap,
                    int method, int is_exit)
{
    double corr = or_fringe_corr(irho, edge, fint, gap, method);
    double fx = irho * tan(edge);
    double fy;
    if (method == 2)
        fy = irho * tan(edge - corr / (1 + r[4])) / (1 + r[4]);
    else if (method == 3)
        fy = is_exit ? irho * tan(edge - corr - r[1] / (1 + r[4]))
                     : irho * tan(edge - corr + r[1] / (1 + r[4]));
    else /* method 1, and the legacy fall-back for 0 and anything else */
        fy = irho * tan(edge - corr / (1 + r[4]));
    r[1] += r[0] * fx;
    r[3] -= r[2] * fy;
}

static void or_rad_edge(double *r, double irho, double edge, double fint, double gap,
                        int method, double sign)
{
    double pn = 1.0 / (1.0 + r[4]);
    double psi;
    if ((fint == 0.0) || (gap == 0.0) || (method == 0))
        psi = 0.0;
    else {
        double s = sin(edge);
        double c = cos(edge);
        psi = irho * gap * fint * (1.0 + s * s) / c;
    }
    double fx = irho * tan(edge);
    double fy;
    if (method == 2)
        fy = irho * tan(edge - psi * pn) * pn;
    else if (method == 3)
        fy = irho * tan(edge - psi + sign * r[1] * pn);
    else
        fy = irho * tan(edge - psi * pn);
    r[1] += r[0] * fx;
    r[3] -= r[2] * fy;
}

/* ---------------------------------------------------------------------- */
/* Quadrupole fringes: quadfringe.c:3-163                                  */

/* Lee-Whiting hard-edge map, sign +1 at entrance, -1 at exit */
static void or_quad_fringe(double *r, double b2, int entrance)
{
    double u = b2 / (12.0 * (1.0 + r[4]));
    double x2 = r[0] * r[0];
    double z2 = r[2] * r[2];
    double xz = r[0] * r[2];
    double gx = u * (x2 + 3 * z2) * r[0];
    double gz = u * (z2 + 3 * x2) * r[2];
    double dpx = 3 * u * (2 * xz * r[3] - (x2 + z2) * r[1]);
    double dpy = 3 * u * (2 * xz * r[1] - (x2 + z2) * r[3]);
    if (entrance) {
        r[0] += gx;
        r[2] -= gz;
        r[5] -= (gz * r[3] - gx * r[1]) / (1 + r[4]);
        r[1] += dpx;
        r[3] -= dpy;
    } else {
        r[0] -= gx;
        r[2] += gz;
        r[5] += (gz * r[3] - gx * r[1]) / (1 + r[4]);
        r[1] -= dpx;
        r[3] += dpy;
    }
}

/* 2x2 blocks of the elegant linear fringe matrix: quadfringe.c:52-93.
 * m = {Rxx, Rxp, Rpx, Rpp, Ryy, Ryp, Rqy, Rqq} */
static void or_elegant_matrix(double m[8], double k1, const double *fi, int part)
{
    double k1sq = k1 * k1;
    double j1x, j2x, j3x, j1y, j2y, j3y;
    if (part == 1) {
        j1x = 1.0 * (k1 * fi[1] - 2 * k1sq * fi[3] / 3. - k1sq * fi[0] * fi[2] / 2);
        j2x = 1.0 * k1 * fi[2];
        j3x = 1.0 * k1sq * (fi[2] + fi[4] + fi[0] * fi[1]);
        j1y = 1.0 * (-k1 * fi[1] - 2 * k1sq * fi[3] / 3. - k1sq * fi[0] * fi[2] / 2);
    } else {
        j1x = 1.0 * (k1 * fi[1] + k1sq * fi[0] * fi[2] / 2);
        j2x = 1.0 * k1 * fi[2];
        j3x = 1.0 * k1sq * (fi[4] - fi[0] * fi[1]);
        j1y = 1.0 * (-k1 * fi[1] + k1sq * fi[0] * fi[2] / 2);
    }
    j2y = -j2x;
    j3y = j3x;
    double ex = exp(j1x);
    double ey = exp(j1y);
    m[0] = ex;
    m[1] = j2x / ex;
    m[2] = ex * j3x;
    m[3] = (1 + j2x * j3x) / ex;
    m[4] = ey;
    m[5] = j2y / ey;
    m[6] = ey * j3y;
    m[7] = (1 + j2y * j3y) / ey;
}

static void or_apply_2x2(double *r, const double m[8], int swap_diag)
{
    double rxx = swap_diag ? m[3] : m[0], rpp = swap_diag ? m[0] : m[3];
    double ryy = swap_diag ? m[7] : m[4], rqq = swap_diag ? m[4] : m[7];
    double x = r[0], px = r[1], y = r[2], py = r[3];
    r[0] = rxx * x + m[1] * px;
    r[1] = m[2] * x + rpp * px;
    r[2] = ryy * y + m[5] * py;
    r[3] = m[6] * y + rqq * py;
}

static void or_quad_fringe_entrance(double *r, const oracle_elem *e, double b2)
{
    if (e->FringeQuadEntrance == 2 && e->fringeIntM0 && e->fringeIntP0) {
        double m[8];
        or_elegant_matrix(m, b2, e->fringeIntP0, 2);
        or_apply_2x2(r, m, 1);
        or_quad_fringe(r, b2, 1);
        or_elegant_matrix(m, b2, e->fringeIntM0, 1);
        or_apply_2x2(r, m, 1);
    } else {
        or_quad_fringe(r, b2, 1);
    }
}

static void or_quad_fringe_exit(double *r, const oracle_elem *e, double b2)
{
    if (e->FringeQuadExit == 2 && e->fringeIntM0 && e->fringeIntP0) {
        double m[8];
        or_elegant_matrix(m, b2, e->fringeIntM0, 1);
        or_apply_2x2(r, m, 0);
        or_quad_fringe(r, b2, 0);
        or_elegant_matrix(m, b2, e->fringeIntP0, 2);
        or_apply_2x2(r, m, 0);
    } else {
        or_quad_fringe(r, b2, 0);
    }
}

/* ---------------------------------------------------------------------- */
/* Per-element state prepared once per call (the reference recomputes the */
/* same element constants inside each pass function).                      */

typedef struct {
    double a[32], b[32];   /* polynomials with the corrector folded in */
    double b2;             /* PolynomB[1] for the quadrupole fringes */
    double l1, l2, k1, k2; /* Forest-Ruth slice lengths */
    double irho;
    double rad_const;
    double inv_scaling;
    double nv, harm;       /* RF */
    int order;
} or_prep;

static int or_prepare(const oracle_elem *e, or_prep *p, double energy, double rest_energy,
                      double T0)
{
    memset(p, 0, sizeof(*p));
    p->inv_scaling = 1.0 / e->FieldScaling;
    switch (e->pass) {
    case OR_STRMPOLESYMPLECTIC4PASS:
    case OR_STRMPOLESYMPLECTIC4RADPASS:
    case OR_BNDMPOLESYMPLECTIC4PASS:
    case OR_BNDMPOLESYMPLECTIC4RADPASS:
    case OR_EXACTMULTIPOLEPASS:
    case OR_EXACTMULTIPOLERADPASS:
    case OR_THINMPOLEPASS: {
        int n = e->MaxOrder;
        if ((e->pass == OR_EXACTMULTIPOLEPASS || e->pass == OR_EXACTMULTIPOLERADPASS) &&
            e->NumIntSteps <= 0)
            return or_fail(3, "NumIntSteps must be positive"); /* ExactMultipolePass.c:126-128 */
        if (n < 0 || n > 30) return or_fail(2, "MaxOrder out of range");
        if (!e->PolynomA || !e->PolynomB) return or_fail(2, "PolynomA/PolynomB missing");
        p->order = n;
        for (int i = 0; i <= n; i++) {
            p->a[i] = e->PolynomA[i];
            p->b[i] = e->PolynomB[i];
        }
        if (e->pass == OR_THINMPOLEPASS) {
            /* ThinMPolePass.c:38-41: angles added without sin or division */
            if (e->KickAngle) {
                p->b[0] -= e->KickAngle[0];
                p->a[0] += e->KickAngle[1];
            }
        } else {
            /* The caller supplies at least two coefficients for the thick
             * magnets, as pyat does (DefaultOrder 1): B[1] drives the
             * quadrupole fringes whatever MaxOrder is. */
            p->b2 = e->PolynomB[1];
            /* StrMPoleSymplectic4Pass.c:39-52 */
            double sl = e->Length / e->NumIntSteps;
            p->l1 = sl * OR_D1;
            p->l2 = sl * OR_D2;
            p->k1 = sl * OR_K1;
            p->k2 = sl * OR_K2;
            if (e->KickAngle) {
                p->b[0] -= sin(e->KickAngle[0]) / e->Length;
                p->a[0] += sin(e->KickAngle[1]) / e->Length;
            }
        }
        if (e->pass == OR_BNDMPOLESYMPLECTIC4PASS || e->pass == OR_BNDMPOLESYMPLECTIC4RADPASS)
            p->irho = e->BendingAngle / e->Length;
        if (e->pass == OR_STRMPOLESYMPLECTIC4RADPASS || e->pass == OR_BNDMPOLESYMPLECTIC4RADPASS ||
            e->pass == OR_EXACTMULTIPOLERADPASS) {
            /* atelem.c:203-211,364-371 */
            if (!(energy > 0.0))
                return or_fail(3, "Energy must be positive. Check lattice or set the track() keyword argument.");
            double gamma = (rest_energy == 0.0) ? 1.0E-9 * energy / OR_E0_GEV : energy / rest_energy;
            p->rad_const = OR_RAD_CONST * pow(gamma, 3);
        }
        break;
    }
    case OR_RFCAVITYPASS:
        if (!(energy > 0.0))
            return or_fail(3, "Energy must be positive. Check lattice or set the track() keyword argument.");
        p->harm = round(e->Frequency * T0); /* RFCavityPass.c:59 */
        p->nv = e->Voltage / energy;        /* RFCavityPass.c:65 */
        break;
    case OR_CAVITYPASS: /* CavityPass.c:66-88: element Energy when the ring's is 0 */
        p->nv = e->Voltage / (energy == 0.0 ? e->Energy : energy);
        break;
    case OR_CORRECTORPASS:
        if (!e->KickAngle) return or_fail(2, "The required attribute KickAngle is missing.");
        break;
    case OR_APERTUREPASS:
        if (!e->Limits) return or_fail(2, "The required attribute Limits is missing.");
        break;
    case OR_IDENTITYPASS:
    case OR_DRIFTPASS:
    case OR_EXACTDRIFTPASS:
        break;
    default:
        return or_fail(4, "PassMethod not in the oracle");
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* One particle through one element.                                        */

static void or_pass(double *r, const oracle_elem *e, const or_prep *p, int nturn, double T0)
{
    if (isnan(r[0])) return; /* lost particles are skipped by every pass */
    const double s = e->FieldScaling;
    switch (e->pass) {
    case OR_IDENTITYPASS: /* IdentityPass.c:20-43 */
        or_entry_transform(r, e);
        or_apertures(r, e);
        or_exit_transform(r, e);
        break;
    case OR_DRIFTPASS: /* DriftPass.c:16-48 */
        or_entry_transform(r, e);
        or_apertures(r, e);
        or_drift6(r, e->Length);
        or_apertures(r, e);
        or_exit_transform(r, e);
        break;
    case OR_APERTUREPASS: /* AperturePass.c:9-26 */
        or_rect_ap(r, e->Limits);
        break;
    case OR_STRMPOLESYMPLECTIC4PASS: /* StrMPoleSymplectic4Pass.c:57-104 */
    case OR_BNDMPOLESYMPLECTIC4PASS: /* BndMPoleSymplectic4Pass.c:75-122 */
    {
        const int bend = (e->pass == OR_BNDMPOLESYMPLECTIC4PASS);
        if (s != 1.0) or_change_pref(r, s);
        double pn = 1.0 / (1.0 + r[4]);
        double nl1 = p->l1 * pn;
        double nl2 = p->l2 * pn;
        or_entry_transform(r, e);
        or_apertures(r, e);
        if (bend)
            or_edge(r, p->irho, e->EntranceAngle, e->FringeInt1, e->FullGap, e->FringeBendEntrance, 0);
        if (e->FringeQuadEntrance && p->b2 != 0) or_quad_fringe_entrance(r, e, p->b2);
        for (int m = 0; m < e->NumIntSteps; m++) {
            if (bend) {
                or_fastdrift(r, nl1);
                or_bnd_kick(r, p->a, p->b, p->k1, p->irho, p->order);
                or_fastdrift(r, nl2);
                or_bnd_kick(r, p->a, p->b, p->k2, p->irho, p->order);
                or_fastdrift(r, nl2);
                or_bnd_kick(r, p->a, p->b, p->k1, p->irho, p->order);
                or_fastdrift(r, nl1);
            } else {
                or_fastdrift(r, nl1);
                or_str_kick(r, p->a, p->b, p->k1, p->order);
                or_fastdrift(r, nl2);
                or_str_kick(r, p->a, p->b, p->k2, p->order);
                or_fastdrift(r, nl2);
                or_str_kick(r, p->a, p->b, p->k1, p->order);
                or_fastdrift(r, nl1);
            }
        }
        if (e->FringeQuadExit && p->b2 != 0) or_quad_fringe_exit(r, e, p->b2);
        if (bend)
            or_edge(r, p->irho, e->ExitAngle, e->FringeInt2, e->FullGap, e->FringeBendExit, 1);
        or_apertures(r, e);
        or_exit_transform(r, e);
        if (s != 1.0) or_change_pref(r, p->inv_scaling);
        break;
    }
    case OR_STRMPOLESYMPLECTIC4RADPASS: /* StrMPoleSymplectic4RadPass.c:61-110 */
    case OR_BNDMPOLESYMPLECTIC4RADPASS: /* BndMPoleSymplectic4RadPass.c:76-123 */
    {
        const int bend = (e->pass == OR_BNDMPOLESYMPLECTIC4RADPASS);
        if (s != 1.0) or_change_pref(r, s);
        or_entry_transform(r, e);
        or_apertures(r, e);
        if (bend)
            or_rad_edge(r, p->irho, e->EntranceAngle, e->FringeInt1, e->FullGap, e->FringeBendEntrance, 1.0);
        if (e->FringeQuadEntrance && p->b2 != 0) or_quad_fringe_entrance(r, e, p->b2);
        for (int m = 0; m < e->NumIntSteps; m++) {
            if (bend) {
                or_rad_drift(r, p->l1);
                or_rad_bnd_kick(r, p->a, p->b, p->order, p->k1, p->irho, p->rad_const);
                or_rad_drift(r, p->l2);
                or_rad_bnd_kick(r, p->a, p->b, p->order, p->k2, p->irho, p->rad_const);
                or_rad_drift(r, p->l2);
                or_rad_bnd_kick(r, p->a, p->b, p->order, p->k1, p->irho, p->rad_const);
                or_rad_drift(r, p->l1);
            } else {
                or_rad_drift(r, p->l1);
                or_rad_str_kick(r, p->a, p->b, p->order, p->k1, p->rad_const);
                or_rad_drift(r, p->l2);
                or_rad_str_kick(r, p->a, p->b, p->order, p->k2, p->rad_const);
                or_rad_drift(r, p->l2);
                or_rad_str_kick(r, p->a, p->b, p->order, p->k1, p->rad_const);
                or_rad_drift(r, p->l1);
            }
        }
        if (e->FringeQuadExit && p->b2 != 0) or_quad_fringe_exit(r, e, p->b2);
        if (bend)
            or_rad_edge(r, p->irho, e->ExitAngle, e->FringeInt2, e->FullGap, e->FringeBendExit, -1.0);
        or_apertures(r, e);
        or_exit_transform(r, e);
        if (s != 1.0) or_change_pref(r, p->inv_scaling);
        break;
    }
    case OR_THINMPOLEPASS: /* ThinMPolePass.c:42-63 */
        if (s != 1.0) or_change_pref(r, s);
        or_entry_transform(r, e);
        or_apertures(r, e);
        or_str_kick(r, p->a, p->b, 1.0, p->order);
        r[1] += e->bax * r[4];
        r[3] -= e->bay * r[4];
        r[5] -= e->bax * r[0] - e->bay * r[2];
        or_exit_transform(r, e);
        if (s != 1.0) or_change_pref(r, p->inv_scaling);
        break;
    case OR_RFCAVITYPASS: /* attrackfunc.c:5-42 */
    {
        const double le = e->Length;
        const double freq = e->Frequency;
        if (le == 0) {
            if (p->nv != 0)
                r[4] += -p->nv * sin(OR_TWOPI * freq * ((r[5] - e->TimeLag) / OR_C0 - (p->harm / freq - T0) * nturn) - e->PhaseLag);
        } else {
            double half = le / 2;
            or_drift6(r, half);
            if (p->nv != 0.0)
                r[4] += -p->nv * sin(OR_TWOPI * freq * ((r[5] - e->TimeLag) / OR_C0 - (p->harm / freq - T0) * nturn) - e->PhaseLag);
            or_drift6(r, half);
        }
        break;
    }
    case OR_EXACTDRIFTPASS: /* ExactDriftPass.c:15-44 */
        or_entry_transform(r, e);
        or_apertures(r, e);
        or_exact_drift(r, e->Length);
        r[5] -= e->Length;
        or_apertures(r, e);
        or_exit_transform(r, e);
        break;
    case OR_EXACTMULTIPOLEPASS:    /* ExactMultipolePass.c:27-96 */
    case OR_EXACTMULTIPOLERADPASS: /* ExactMultipoleRadPass.c:26-104 */
    {
        const int rad = (e->pass == OR_EXACTMULTIPOLERADPASS);
        if (s != 1.0) or_change_pref(r, s);
        or_entry_transform(r, e);
        or_apertures(r, e);
        if (e->FringeQuadEntrance) or_multipole_fringe(r, p->a, p->b, p->order, 1.0);
        for (int m = 0; m < e->NumIntSteps; m++) {
            if (rad) {
                or_exact_rad_drift(r, p->l1);
                or_exact_rad_kick(r, p->a, p->b, p->order, p->k1, p->rad_const);
                or_exact_rad_drift(r, p->l2);
                or_exact_rad_kick(r, p->a, p->b, p->order, p->k2, p->rad_const);
                or_exact_rad_drift(r, p->l2);
                or_exact_rad_kick(r, p->a, p->b, p->order, p->k1, p->rad_const);
                or_exact_rad_drift(r, p->l1);
            } else {
                or_exact_drift(r, p->l1);
                or_str_kick(r, p->a, p->b, p->k1, p->order);
                or_exact_drift(r, p->l2);
                or_str_kick(r, p->a, p->b, p->k2, p->order);
                or_exact_drift(r, p->l2);
                or_str_kick(r, p->a, p->b, p->k1, p->order);
                or_exact_drift(r, p->l1);
            }
        }
        r[5] -= e->Length;
        if (e->FringeQuadExit) or_multipole_fringe(r, p->a, p->b, p->order, -1.0);
        or_apertures(r, e);
        or_exit_transform(r, e);
        if (s != 1.0) or_change_pref(r, p->inv_scaling);
        break;
    }
    case OR_CAVITYPASS: /* CavityPass.c:18-56 */
    {
        const double le = e->Length;
        if (le == 0) {
            r[4] += -p->nv * sin(OR_TWOPI * e->Frequency * (r[5] - e->TimeLag) / OR_C0 - e->PhaseLag);
        } else {
            double half = le / 2;
            or_drift6(r, half);
            r[4] += -p->nv * sin(OR_TWOPI * e->Frequency * (r[5] - e->TimeLag) / OR_C0 - e->PhaseLag);
            or_drift6(r, half);
        }
        break;
    }
    case OR_CORRECTORPASS: /* CorrectorPass.c:25-90 */
    {
        const double xk = e->KickAngle[0], yk = e->KickAngle[1], le = e->Length;
        if (s != 1.0) or_change_pref(r, s);
        if (le == 0.0) {
            or_entry_transform(r, e);
            or_apertures(r, e);
            r[1] += xk;
            r[3] += yk;
        } else {
            double pn = 1.0 / (1.0 + r[4]);
            double nl = le * pn;
            or_entry_transform(r, e);
            or_apertures(r, e);
            r[5] += nl * pn * (xk * xk / 3 + yk * yk / 3 + r[1] * r[1] + r[3] * r[3] + r[1] * xk + r[3] * yk) / 2;
            r[0] += nl * (r[1] + xk / 2);
            r[1] += xk;
            r[2] += nl * (r[3] + yk / 2);
            r[3] += yk;
            or_apertures(r, e);
        }
        or_exit_transform(r, e);
        if (s != 1.0) or_change_pref(r, p->inv_scaling);
        break;
    }
    default:
        break;
    }
}

/* at.c:115-139 loss predicate */
static int or_is_lost(const double *r)
{
    for (int n = 0; n < 6; n++)
        if (!isfinite(r[n]) || ((fabs(r[n]) > 1) && n < 5)) return 1;
    return 0;
}

static void or_kill(double *r)
{
    r[0] = NAN;
    r[1] = r[2] = r[3] = r[4] = r[5] = 0;
}

/* Refpt slot assignment following at.c:618-651: the output pointer advances
 * once per copy, so the slot of a copy is its rank in the copy sequence. */
static int or_turn_slots(const uint32_t *refpts, int nref, int start, int end, int nelem,
                         int *slot /* nelem+1 entries */)
{
    for (int e = 0; e <= nelem; e++) slot[e] = -1;
    int idx = 0;
    uint32_t next = (idx < nref) ? refpts[idx++] : (uint32_t)INT_MAX;
    int k = 0;
    for (int e = start; e < end; e++) {
        if ((uint32_t)e == next) {
            slot[e] = k++;
            next = (idx < nref) ? refpts[idx++] : (uint32_t)INT_MAX;
        }
    }
    if ((uint32_t)nelem == next) slot[nelem] = k++;
    return k;
}

int oracle_atpass(const oracle_elem *elems, int32_t nelem, double *r_in,
                  int32_t num_particles, int32_t nturns,
                  const uint32_t *refpts, int32_t num_refpts, int32_t turn,
                  double energy, double rest_energy, int32_t losses,
                  int32_t start_elem, int32_t end_elem, double *r_out,
                  uint8_t *lost_flag, uint32_t *lost_turn, uint32_t *lost_elem,
                  double *lost_coord, int32_t nthreads)
{
    or_errbuf[0] = 0;
    /* at.c:514-536 partial-turn checks */
    if (end_elem == 0) end_elem = nelem;
    if (nelem != 0) {
        if (start_elem < 0 || start_elem >= nelem) return or_fail(1, "start_elem out of range");
        if (end_elem < 0 || end_elem > nelem) return or_fail(1, "end_elem out of range");
        if (start_elem > end_elem) return or_fail(1, "end_elem must be greater that start_elem");
        if ((end_elem != nelem) && (nturns != 1))
            return or_fail(1, "partial turn tracking not ending at the end of the ring requires nturns=1");
    } else {
        start_elem = 0;
        end_elem = 0;
    }
    double ring_length = 0.0;
    for (int e = 0; e < nelem; e++) ring_length += elems[e].Length;
    const double T0 = ring_length / OR_C0; /* at.c:540-551 */

    or_prep *prep = (or_prep *)calloc(nelem > 0 ? nelem : 1, sizeof(or_prep));
    int *slot_first = (int *)malloc(sizeof(int) * (nelem + 1));
    int *slot_std = (int *)malloc(sizeof(int) * (nelem + 1));
    int rc = 0;
    for (int e = 0; e < nelem && rc == 0; e++) rc = or_prepare(&elems[e], &prep[e], energy, rest_energy, T0);
    if (rc) {
        free(prep); free(slot_first); free(slot_std);
        return rc;
    }
    int n_first = or_turn_slots(refpts, num_refpts, start_elem, end_elem, nelem, slot_first);
    int n_std = or_turn_slots(refpts, num_refpts, 0, end_elem, nelem, slot_std);
    const long long np6 = 6LL * num_particles;

#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int c = 0; c < num_particles; c++) {
        double r[6];
        memcpy(r, r_in + 6LL * c, sizeof(r));
        int lost = 0;
        if (losses) {
            lost_flag[c] = 0;
            lost_turn[c] = 0;
            lost_elem[c] = 0;
            memset(lost_coord + 6LL * c, 0, 6 * sizeof(double));
        }
        for (int t = 0; t < nturns; t++) {
            const int nturn = turn + t;
            const int e0 = (t == 0) ? start_elem : 0;
            const int *slot = (t == 0) ? slot_first : slot_std;
            const long long base = (t == 0) ? 0 : (long long)n_first + (long long)(t - 1) * n_std;
            for (int e = e0; e < end_elem; e++) {
                if (slot[e] >= 0) memcpy(r_out + (base + slot[e]) * np6 + 6LL * c, r, sizeof(r));
                or_pass(r, &elems[e], &prep[e], nturn, T0);
                if (losses) {
                    if (!lost && or_is_lost(r)) {
                        lost = 1;
                        lost_flag[c] = 1;
                        lost_turn[c] = (uint32_t)nturn;
                        lost_elem[c] = (uint32_t)e;
                        memcpy(lost_coord + 6LL * c, r, sizeof(r));
                        or_kill(r);
                    }
                } else if (isfinite(r[0]) && or_is_lost(r)) {
                    or_kill(r); /* at.c:142-161 */
                }
            }
            if (slot[nelem] >= 0) memcpy(r_out + (base + slot[nelem]) * np6 + 6LL * c, r, sizeof(r));
        }
        memcpy(r_in + 6LL * c, r, sizeof(r));
    }
    free(prep);
    free(slot_first);
    free(slot_std);
    return 0;
}

int oracle_elempass(const oracle_elem *elem, double *r_in, int32_t num_particles,
                    double energy, double rest_energy)
{
    or_errbuf[0] = 0;
    or_prep p;
    int rc = or_prepare(elem, &p, energy, rest_energy, 0.0);
    if (rc) return rc;
    for (int c = 0; c < num_particles; c++) or_pass(r_in + 6LL * c, elem, &p, 0, 0.0);
    return 0;
}
