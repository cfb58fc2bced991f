/*
 * ref_driver.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Runs the reference's own integrator cores, compiled in place from
 * /root/reference/atintegrators/<Pass>.c into oracle/_ref/lib<Pass>.so (one
 * shared object per pass method: atconstants.h:10-22 defines non-static
 * globals that would collide, which is also why pyat builds one extension
 * per integrator, setup.py:192-203).
 *
 * The reference's trackFunction entry points need the Python C-API and the
 * pyat package (atelem.c:192-354), which cannot run here, so this driver
 * restates what trackFunction and at.c do around the cores:
 *   - parameter defaults and derived values (irho = BendingAngle/Length,
 *     gamma = atGamma(), HarmNumber = round(Frequency*T0), nv = V/E);
 *   - the element-outer turn loop of at.c:596-654 with the refpt copies and
 *     the serial checkiflost()/setlost() scan after every element.
 * Structure and OpenMP usage match the reference CPU path, so this is also
 * the timed "reference CPU" baseline.
 */
#include "oracle.h"

#include <dlfcn.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef void (*fn_drift)(double *, double, const double *, const double *, const double *,
                         const double *, double *, double *, int);
typedef void (*fn_identity)(double *, const double *, const double *, const double *,
                            const double *, const double *, const double *, int);
typedef void (*fn_str)(double *, double, double *, double *, int, int, int, int, double *,
                       double *, double *, double *, double *, double *, double *, double *,
                       double *, double, int);
typedef void (*fn_bnd)(double *, double, double, double *, double *, int, int, double, double,
                       int, int, double, double, double, int, int, double *, double *,
                       double *, double *, double *, double *, double *, double *, double *,
                       double, int);
typedef void (*fn_strrad)(double *, double, double *, double *, int, int, int, int, double *,
                          double *, double *, double *, double *, double *, double *, double *,
                          double *, double, double, int, double *);
typedef void (*fn_bndrad)(double *, double, double, double *, double *, int, int, double,
                          double, int, int, double, double, double, int, int, double *,
                          double *, double *, double *, double *, double *, double *,
                          double *, double *, double, double, int, double *);
typedef void (*fn_thin)(double *, double *, double *, int, double, double, double *, double *,
                        double *, double *, double *, double *, double *, double, int);
typedef void (*fn_rf)(double *, double, double, double, double, double, double, int, double,
                      int);
typedef void (*fn_corr)(double *, double, double, double, double *, double *, double *,
                        double *, double *, double *, double, int);
typedef void (*fn_ap)(double *, double *, int);
typedef void (*fn_cav)(double *, double, double, double, double, double, int);
typedef void (*fn_exdrift)(double *, double, const double *, const double *, const double *,
                           const double *, double *, double *, int);
typedef void (*fn_exmpole)(double *, double, double *, double *, int, int, int, int, double *,
                           double *, double *, double *, double *, double *, double *, double,
                           int);
typedef void (*fn_exmpolerad)(double *, double, double *, double *, int, int, int, int,
                              double *, double *, double *, double *, double *, double *,
                              double *, double, double, int, double *);

static char rf_err[1024];
static char rf_dir[400] = "";
static void *rf_sym[16];

const char *ref_last_error(void) { return rf_err; }

static int rf_fail(int code, const char *msg)
{
    snprintf(rf_err, sizeof(rf_err), "%s", msg);
    return code;
}

int ref_set_library_dir(const char *dir)
{
    snprintf(rf_dir, sizeof(rf_dir), "%s", dir);
    memset(rf_sym, 0, sizeof(rf_sym));
    return 0;
}

static const char *rf_name(int pass)
{
    switch (pass) {
    case OR_IDENTITYPASS: return "IdentityPass";
    case OR_DRIFTPASS: return "DriftPass";
    case OR_BNDMPOLESYMPLECTIC4PASS: return "BndMPoleSymplectic4Pass";
    case OR_BNDMPOLESYMPLECTIC4RADPASS: return "BndMPoleSymplectic4RadPass";
    case OR_STRMPOLESYMPLECTIC4PASS: return "StrMPoleSymplectic4Pass";
    case OR_STRMPOLESYMPLECTIC4RADPASS: return "StrMPoleSymplectic4RadPass";
    case OR_RFCAVITYPASS: return "RFCavityPass";
    case OR_THINMPOLEPASS: return "ThinMPolePass";
    case OR_CORRECTORPASS: return "CorrectorPass";
    case OR_APERTUREPASS: return "AperturePass";
    case OR_CAVITYPASS: return "CavityPass";
    case OR_EXACTDRIFTPASS: return "ExactDriftPass";
    case OR_EXACTMULTIPOLEPASS: return "ExactMultipolePass";
    case OR_EXACTMULTIPOLERADPASS: return "ExactMultipoleRadPass";
    default: return NULL;
    }
}

/* Resolve the core function of one pass method (cf. at.c:236-273). */
static void *rf_resolve(int pass)
{
    if (pass < 0 || pass >= 16) return NULL;
    if (rf_sym[pass]) return rf_sym[pass];
    const char *name = rf_name(pass);
    if (!name) return NULL;
    char path[480];
    snprintf(path, sizeof(path), "%s/lib%s.so", rf_dir, name);
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        snprintf(rf_err, sizeof(rf_err), "dlopen %s: %s", path, dlerror());
        return NULL;
    }
    void *f = dlsym(h, name);
    if (!f) snprintf(rf_err, sizeof(rf_err), "dlsym %s failed", name);
    rf_sym[pass] = f;
    return f;
}

#define RF_E0_GEV 510.99895069e-6 /* atconstants.h:27 */
#define RF_C0 2.99792458e8

typedef struct {
    double a[32], b[32]; /* writable copies: the cores fold KickAngle into them */
    double gamma, irho, harm, nv;
} rf_prep;

/* Restates the derived values of the trackFunction wrappers. */
static int rf_prepare(const oracle_elem *e, rf_prep *p, double energy, double rest_energy,
                      double T0)
{
    memset(p, 0, sizeof(*p));
    if (e->PolynomA && e->PolynomB) {
        int n = e->MaxOrder;
        int ncopy = (n + 1 > 2) ? n + 1 : 2;
        if (n < 0 || n > 30) return rf_fail(2, "MaxOrder out of range");
        for (int i = 0; i < ncopy; i++) {
            p->a[i] = e->PolynomA[i];
            p->b[i] = e->PolynomB[i];
        }
    }
    if (e->pass == OR_BNDMPOLESYMPLECTIC4PASS || e->pass == OR_BNDMPOLESYMPLECTIC4RADPASS)
        p->irho = e->BendingAngle / e->Length;
    if (e->pass == OR_BNDMPOLESYMPLECTIC4RADPASS || e->pass == OR_STRMPOLESYMPLECTIC4RADPASS ||
        e->pass == OR_EXACTMULTIPOLERADPASS) {
        if (!(energy > 0.0)) return rf_fail(3, "Energy must be positive");
        p->gamma = (rest_energy == 0.0) ? 1.0E-9 * energy / RF_E0_GEV : energy / rest_energy;
    }
    if (e->pass == OR_CAVITYPASS) /* CavityPass.c:84-86 */
        p->nv = e->Voltage / (energy == 0.0 ? e->Energy : energy);
    if ((e->pass == OR_EXACTMULTIPOLEPASS || e->pass == OR_EXACTMULTIPOLERADPASS) &&
        e->NumIntSteps <= 0)
        return rf_fail(3, "NumIntSteps must be positive");
    if (e->pass == OR_RFCAVITYPASS) {
        if (!(energy > 0.0)) return rf_fail(3, "Energy must be positive");
        p->harm = round(e->Frequency * T0);
        p->nv = e->Voltage / energy;
    }
    return 0;
}

#define D(x) ((double *)(x))

static int rf_call(const oracle_elem *e, rf_prep *p, double *r, int np, int nturn, double T0)
{
    void *f = rf_resolve(e->pass);
    if (!f) return rf_fail(4, rf_err[0] ? rf_err : "PassMethod not available");
    switch (e->pass) {
    case OR_IDENTITYPASS:
        ((fn_identity)f)(r, e->T1, e->T2, e->R1, e->R2, e->RApertures, e->EApertures, np);
        break;
    case OR_DRIFTPASS:
        ((fn_drift)f)(r, e->Length, e->T1, e->T2, e->R1, e->R2, D(e->RApertures),
                      D(e->EApertures), np);
        break;
    case OR_STRMPOLESYMPLECTIC4PASS:
        ((fn_str)f)(r, e->Length, p->a, p->b, e->MaxOrder, e->NumIntSteps,
                    e->FringeQuadEntrance, e->FringeQuadExit, D(e->fringeIntM0),
                    D(e->fringeIntP0), D(e->T1), D(e->T2), D(e->R1), D(e->R2),
                    D(e->RApertures), D(e->EApertures), D(e->KickAngle), e->FieldScaling, np);
        break;
    case OR_BNDMPOLESYMPLECTIC4PASS:
        ((fn_bnd)f)(r, e->Length, p->irho, p->a, p->b, e->MaxOrder, e->NumIntSteps,
                    e->EntranceAngle, e->ExitAngle, e->FringeBendEntrance, e->FringeBendExit,
                    e->FringeInt1, e->FringeInt2, e->FullGap, e->FringeQuadEntrance,
                    e->FringeQuadExit, D(e->fringeIntM0), D(e->fringeIntP0), D(e->T1),
                    D(e->T2), D(e->R1), D(e->R2), D(e->RApertures), D(e->EApertures),
                    D(e->KickAngle), e->FieldScaling, np);
        break;
    case OR_STRMPOLESYMPLECTIC4RADPASS:
        ((fn_strrad)f)(r, e->Length, p->a, p->b, e->MaxOrder, e->NumIntSteps,
                       e->FringeQuadEntrance, e->FringeQuadExit, D(e->fringeIntM0),
                       D(e->fringeIntP0), D(e->T1), D(e->T2), D(e->R1), D(e->R2),
                       D(e->RApertures), D(e->EApertures), D(e->KickAngle), e->FieldScaling,
                       p->gamma, np, NULL);
        break;
    case OR_BNDMPOLESYMPLECTIC4RADPASS:
        ((fn_bndrad)f)(r, e->Length, p->irho, p->a, p->b, e->MaxOrder, e->NumIntSteps,
                       e->EntranceAngle, e->ExitAngle, e->FringeBendEntrance,
                       e->FringeBendExit, e->FringeInt1, e->FringeInt2, e->FullGap,
                       e->FringeQuadEntrance, e->FringeQuadExit, D(e->fringeIntM0),
                       D(e->fringeIntP0), D(e->T1), D(e->T2), D(e->R1), D(e->R2),
                       D(e->RApertures), D(e->EApertures), D(e->KickAngle), e->FieldScaling,
                       p->gamma, np, NULL);
        break;
    case OR_THINMPOLEPASS:
        ((fn_thin)f)(r, p->a, p->b, e->MaxOrder, e->bax, e->bay, D(e->T1), D(e->T2),
                     D(e->R1), D(e->R2), D(e->RApertures), D(e->EApertures),
                     D(e->KickAngle), e->FieldScaling, np);
        break;
    case OR_RFCAVITYPASS:
        ((fn_rf)f)(r, e->Length, p->nv, e->Frequency, p->harm, e->TimeLag, e->PhaseLag, nturn,
                   T0, np);
        break;
    case OR_CORRECTORPASS:
        if (!e->KickAngle) return rf_fail(2, "KickAngle missing");
        ((fn_corr)f)(r, e->KickAngle[0], e->KickAngle[1], e->Length, D(e->T1), D(e->T2),
                     D(e->R1), D(e->R2), D(e->RApertures), D(e->EApertures), e->FieldScaling,
                     np);
        break;
    case OR_APERTUREPASS:
        if (!e->Limits) return rf_fail(2, "Limits missing");
        ((fn_ap)f)(r, D(e->Limits), np);
        break;
    case OR_CAVITYPASS:
        ((fn_cav)f)(r, e->Length, p->nv, e->Frequency, e->TimeLag, e->PhaseLag, np);
        break;
    case OR_EXACTDRIFTPASS:
        ((fn_exdrift)f)(r, e->Length, e->T1, e->T2, e->R1, e->R2, D(e->RApertures),
                        D(e->EApertures), np);
        break;
    case OR_EXACTMULTIPOLEPASS:
        ((fn_exmpole)f)(r, e->Length, p->a, p->b, e->MaxOrder, e->NumIntSteps,
                        e->FringeQuadEntrance, e->FringeQuadExit, D(e->T1), D(e->T2), D(e->R1),
                        D(e->R2), D(e->RApertures), D(e->EApertures), D(e->KickAngle),
                        e->FieldScaling, np);
        break;
    case OR_EXACTMULTIPOLERADPASS:
        ((fn_exmpolerad)f)(r, e->Length, p->a, p->b, e->MaxOrder, e->NumIntSteps,
                           e->FringeQuadEntrance, e->FringeQuadExit, D(e->T1), D(e->T2),
                           D(e->R1), D(e->R2), D(e->RApertures), D(e->EApertures),
                           D(e->KickAngle), e->FieldScaling, p->gamma, np, NULL);
        break;
    default:
        return rf_fail(4, "PassMethod not available");
    }
    return 0;
}

/* at.c:115-139 (serial, as in the reference) */
static void rf_checkiflost(double *r, int np, int elem, int turn, uint32_t *xt, uint32_t *xe,
                           uint8_t *xl, double *xc)
{
    for (int c = 0; c < np; c++) {
        if (xl[c]) continue;
        double *r6 = r + 6LL * c;
        for (int n = 0; n < 6; n++) {
            if (!isfinite(r6[n]) || ((fabs(r6[n]) > 1) && n < 5)) {
                xl[c] = 1;
                xt[c] = (uint32_t)turn;
                xe[c] = (uint32_t)elem;
                memcpy(xc + 6LL * c, r6, 6 * sizeof(double));
                r6[0] = NAN;
                r6[1] = r6[2] = r6[3] = r6[4] = r6[5] = 0;
                break;
            }
        }
    }
}

/* at.c:142-161 */
static void rf_setlost(double *r, int np)
{
    for (int c = 0; c < np; c++) {
        double *r6 = r + 6LL * c;
        if (!isfinite(r6[0])) continue;
        for (int n = 0; n < 6; n++) {
            if (!isfinite(r6[n]) || ((fabs(r6[n]) > 1) && n < 5)) {
                r6[0] = NAN;
                r6[1] = r6[2] = r6[3] = r6[4] = r6[5] = 0;
                break;
            }
        }
    }
}

int ref_atpass(const oracle_elem *elems, int32_t nelem, double *r_in, int32_t np,
               int32_t nturns, const uint32_t *refpts, int32_t nref, int32_t turn,
               double energy, double rest_energy, int32_t losses, int32_t start_elem,
               int32_t end_elem, double *r_out, uint8_t *lost_flag, uint32_t *lost_turn,
               uint32_t *lost_elem, double *lost_coord, int32_t nthreads)
{
    rf_err[0] = 0;
    if (end_elem == 0) end_elem = nelem;
    if (nelem != 0) {
        if (start_elem < 0 || start_elem >= nelem) return rf_fail(1, "start_elem out of range");
        if (end_elem < 0 || end_elem > nelem) return rf_fail(1, "end_elem out of range");
        if (start_elem > end_elem) return rf_fail(1, "end_elem must be greater that start_elem");
        if ((end_elem != nelem) && (nturns != 1)) return rf_fail(1, "partial turn requires nturns=1");
    } else {
        start_elem = 0;
        end_elem = 0;
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    double ring_length = 0.0;
    for (int e = 0; e < nelem; e++) ring_length += elems[e].Length;
    const double T0 = ring_length / RF_C0;
    rf_prep *prep = (rf_prep *)calloc(nelem > 0 ? nelem : 1, sizeof(rf_prep));
    for (int e = 0; e < nelem; e++) {
        int rc = rf_prepare(&elems[e], &prep[e], energy, rest_energy, T0);
        if (rc) {
            free(prep);
            return rc;
        }
    }
    if (losses) {
        memset(lost_flag, 0, np);
        memset(lost_turn, 0, sizeof(uint32_t) * np);
        memset(lost_elem, 0, sizeof(uint32_t) * np);
        memset(lost_coord, 0, sizeof(double) * 6 * np);
    }
    const size_t np6 = 6 * (size_t)np;
    double *drout = r_out;
    int nturn = turn;
    for (int t = 0; t < nturns; t++) {
        int idx = 0;
        uint32_t next = (idx < nref) ? refpts[idx++] : (uint32_t)INT_MAX;
        for (int e = start_elem; e < end_elem; e++) {
            if ((uint32_t)e == next) {
                memcpy(drout, r_in, np6 * sizeof(double));
                drout += np6;
                next = (idx < nref) ? refpts[idx++] : (uint32_t)INT_MAX;
            }
            int rc = rf_call(&elems[e], &prep[e], r_in, np, nturn, T0);
            if (rc) {
                free(prep);
                return rc;
            }
            if (losses)
                rf_checkiflost(r_in, np, e, nturn, lost_turn, lost_elem, lost_flag, lost_coord);
            else
                rf_setlost(r_in, np);
        }
        if ((uint32_t)nelem == next) {
            memcpy(drout, r_in, np6 * sizeof(double));
            drout += np6;
        }
        nturn++;
        start_elem = 0;
    }
    free(prep);
    return 0;
}

/*
 * Tangent map of one multipole element at one particle, from the reference
 * itself: the radiating cores (StrMPoleSymplectic4RadPass.c:83-91,
 * BndMPoleSymplectic4RadPass.c:86-108), given a diffusion-matrix pointer,
 * apply B <- M B M^T with the per-step Jacobians M of diff_drift.c,
 * diff_str_kick.c / diff_bnd_kick.c (thinkickM) and diff_bend_fringe.c.
 * With gamma = 1e-9 the radiation and diffusion terms are below FP64
 * resolution, so B_out = M B_in M^T with M the Jacobian of the
 * non-radiating body. Starting from B_in = e_j e_4^T gives
 * B_out = (M e_j)(M e_4)^T, and since M[4][4] = 1 the column 4 of B_out
 * is column j of M. Six runs fill m66 (row-major). T1/T2, R1/R2 and the
 * quadrupole fringes are not propagated by the cores: use elements
 * without them. r6 is left unchanged.
 */
int ref_tangent(const oracle_elem *e, const double *r6, double *m66)
{
    rf_err[0] = 0;
    oracle_elem rad = *e;
    if (e->pass == OR_STRMPOLESYMPLECTIC4PASS || e->pass == OR_STRMPOLESYMPLECTIC4RADPASS)
        rad.pass = OR_STRMPOLESYMPLECTIC4RADPASS;
    else if (e->pass == OR_BNDMPOLESYMPLECTIC4PASS || e->pass == OR_BNDMPOLESYMPLECTIC4RADPASS)
        rad.pass = OR_BNDMPOLESYMPLECTIC4RADPASS;
    else
        return rf_fail(4, "tangent maps exist for the multipole passes only");
    void *f = rf_resolve(rad.pass);
    if (!f) return rf_fail(4, rf_err);
    for (int j = 0; j < 6; j++) {
        rf_prep p;
        int rc = rf_prepare(&rad, &p, 1.0e-9, 1.0, 0.0); /* gamma = 1e-9 */
        if (rc) return rc;
        double r[6], bdiff[36];
        memcpy(r, r6, sizeof(r));
        memset(bdiff, 0, sizeof(bdiff));
        bdiff[j + 6 * 4] = 1.0;
        if (rad.pass == OR_STRMPOLESYMPLECTIC4RADPASS)
            ((fn_strrad)f)(r, e->Length, p.a, p.b, e->MaxOrder, e->NumIntSteps,
                           e->FringeQuadEntrance, e->FringeQuadExit, D(e->fringeIntM0),
                           D(e->fringeIntP0), D(e->T1), D(e->T2), D(e->R1), D(e->R2),
                           D(e->RApertures), D(e->EApertures), D(e->KickAngle), e->FieldScaling,
                           p.gamma, 1, bdiff);
        else
            ((fn_bndrad)f)(r, e->Length, p.irho, p.a, p.b, e->MaxOrder, e->NumIntSteps,
                           e->EntranceAngle, e->ExitAngle, e->FringeBendEntrance,
                           e->FringeBendExit, e->FringeInt1, e->FringeInt2, e->FullGap,
                           e->FringeQuadEntrance, e->FringeQuadExit, D(e->fringeIntM0),
                           D(e->fringeIntP0), D(e->T1), D(e->T2), D(e->R1), D(e->R2),
                           D(e->RApertures), D(e->EApertures), D(e->KickAngle), e->FieldScaling,
                           p.gamma, 1, bdiff);
        for (int i = 0; i < 6; i++) m66[6 * i + j] = bdiff[i + 6 * 4];
    }
    return 0;
}

/*
 * Diffusion matrix of one element from the orbit at its entrance, as
 * at_diffmatrix (pyat/at.c:755-831) computes it: the element's own pass
 * (the radiating cores propagate bdiff; the others leave it at zero) on one
 * particle with the ring energy and rest energy. bdiff: 36 doubles,
 * column-major.
 */
int ref_diffusion(const oracle_elem *e, const double *orbit, double *bdiff, double energy,
                  double rest_energy)
{
    rf_err[0] = 0;
    for (int i = 0; i < 36; i++) bdiff[i] = 0.0;
    if (e->pass != OR_STRMPOLESYMPLECTIC4RADPASS && e->pass != OR_BNDMPOLESYMPLECTIC4RADPASS &&
        e->pass != OR_EXACTMULTIPOLERADPASS)
        return 0;
    rf_prep p;
    int rc = rf_prepare(e, &p, energy, rest_energy, 0.0);
    if (rc) return rc;
    void *f = rf_resolve(e->pass);
    if (!f) return rf_fail(4, rf_err);
    double r[6];
    memcpy(r, orbit, sizeof(r));
    if (e->pass == OR_STRMPOLESYMPLECTIC4RADPASS)
        ((fn_strrad)f)(r, e->Length, p.a, p.b, e->MaxOrder, e->NumIntSteps, e->FringeQuadEntrance,
                       e->FringeQuadExit, D(e->fringeIntM0), D(e->fringeIntP0), D(e->T1), D(e->T2),
                       D(e->R1), D(e->R2), D(e->RApertures), D(e->EApertures), D(e->KickAngle),
                       e->FieldScaling, p.gamma, 1, bdiff);
    else if (e->pass == OR_BNDMPOLESYMPLECTIC4RADPASS)
        ((fn_bndrad)f)(r, e->Length, p.irho, p.a, p.b, e->MaxOrder, e->NumIntSteps,
                       e->EntranceAngle, e->ExitAngle, e->FringeBendEntrance, e->FringeBendExit,
                       e->FringeInt1, e->FringeInt2, e->FullGap, e->FringeQuadEntrance,
                       e->FringeQuadExit, D(e->fringeIntM0), D(e->fringeIntP0), D(e->T1), D(e->T2),
                       D(e->R1), D(e->R2), D(e->RApertures), D(e->EApertures), D(e->KickAngle),
                       e->FieldScaling, p.gamma, 1, bdiff);
    else
        ((fn_exmpolerad)f)(r, e->Length, p.a, p.b, e->MaxOrder, e->NumIntSteps,
                           e->FringeQuadEntrance, e->FringeQuadExit, D(e->T1), D(e->T2),
                           D(e->R1), D(e->R2), D(e->RApertures), D(e->EApertures),
                           D(e->KickAngle), e->FieldScaling, p.gamma, 1, bdiff);
    return 0;
}

int ref_elempass(const oracle_elem *e, double *r_in, int32_t np, double energy,
                 double rest_energy)
{
    rf_err[0] = 0;
    rf_prep p;
    int rc = rf_prepare(e, &p, energy, rest_energy, 0.0);
    if (rc) return rc;
    return rf_call(e, &p, r_in, np, 0, 0.0);
}
