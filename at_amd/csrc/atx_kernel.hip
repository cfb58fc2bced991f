// atx_kernel.hip -- persistent FP64 tracking kernel for gfx950 (MI355X).
//
// One lane tracks one particle through every element of every turn in a
// single launch; the particle's 6-vector lives in registers from the first
// load to the final store, and HBM is touched only at start, at refpt
// copies, on a loss and at the end. Element parameters are compile-time
// prepared by atx_host.cpp and read through the constant address space, so
// the uniform per-element values land in SGPRs via scalar loads (every lane
// of a workgroup walks the same element sequence; no divergence on the
// element switch).
//
// This file is compiled twice:
//   ATX_STRICT=1, -ffp-contract=off: the reference operation order of
//     atintegrators/*.c, with the same divisions, so results agree bit for
//     bit with the CPU reference except where a device transcendental (sin in
//     RFCavityPass, tan in per-particle edge fringes) enters;
//   ATX_STRICT=0, -ffp-contract=fast: the same maps with FMA contraction,
//     hand-fused Horner steps and the 1/(1+delta) reciprocals hoisted where
//     delta is constant; agreement with the reference is to rounding.
#include <hip/hip_runtime.h>

#include "atx_device.h"
#include "atx_libm.h"

#ifndef ATX_STRICT
#error "ATX_STRICT must be defined to 0 or 1"
#endif
// One object per launch shape (particles per lane ATX_PPT, slices per loop
// iteration ATX_UNROLL) and build, compiled in parallel; the helper kernels
// (repack, live map, diffusion matrices) go into the strict 1x1 object.
#if !defined(ATX_PPT) || !defined(ATX_UNROLL)
#error "ATX_PPT (1, 2) and ATX_UNROLL (1, 2, 4) must be defined"
#endif
#define ATX_AUX (ATX_STRICT && ATX_PPT == 1 && ATX_UNROLL == 1)
#define ATX_PASTE6(a, b, c, d, e, f) a##b##c##d##e##f
#define ATX_NAME(build, p, u) ATX_PASTE6(atx_launch_track_, build, _p, p, _u, u)

#if ATX_STRICT
#define ATX_NS strict_ns
#define ATX_LAUNCH ATX_NAME(strict, ATX_PPT, ATX_UNROLL)
#else
#define ATX_NS fast_ns
#define ATX_LAUNCH ATX_NAME(fast, ATX_PPT, ATX_UNROLL)
#endif

namespace atx {
namespace ATX_NS {

// Read-only element data: constant address space -> scalar loads.
typedef __attribute__((address_space(4))) const double cdouble;
typedef __attribute__((address_space(4))) const DevElem cDevElem;
typedef __attribute__((address_space(4))) const Instance cInstance;
typedef __attribute__((address_space(4))) const int32_t cint;

struct R6 {
  double x, px, y, py, dp, ct;
};

// Forest-Ruth 4th-order coefficients (atintegrators/atconstants.h:15-19);
// only their products with the slice length are used, prepared on the host.

__device__ __forceinline__ void add_t(R6& r, cdouble* t) {
  r.x += t[0];
  r.px += t[1];
  r.y += t[2];
  r.py += t[3];
  r.dp += t[4];
  r.ct += t[5];
}

// Per-element (not per-slice) maps that are large or use transcendentals are
// kept out of line: inlined into every element variant they would multiply
// the kernel's code size and register demand (a 6x6 matrix alone holds 72
// SGPRs), while they run at most a few times per element.

// r <- M r with M column-major (atlalib.c:17-32); the sum starts from 0.0 as
// in the reference.
__device__ __noinline__ R6 mul_r_v(R6 r, cdouble* m) {
  const double v[6] = {r.x, r.px, r.y, r.py, r.dp, r.ct};
  double o[6];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) s += m[i + 6 * j] * v[j];
    o[i] = s;
  }
  return R6{o[0], o[1], o[2], o[3], o[4], o[5]};
}

__device__ __forceinline__ void mul_r(R6& r, cdouble* m) { r = mul_r_v(r, m); }

// atlalib.c:70-75
__device__ __noinline__ R6 change_pref_v(R6 r, double s) {
  r.px /= s;
  r.py /= s;
  r.dp = (r.dp + 1.0 - s) / s;
  return r;
}

__device__ __forceinline__ void change_pref(R6& r, double s) { r = change_pref_v(r, s); }

// Apertures are decisions: evaluated without contraction in both builds so
// that the loss element is the reference's (atlalib.c:164-190).
__device__ __forceinline__ void rect_ap(R6& r, cdouble* lim) {
#pragma clang fp contract(off)
  if (r.x < lim[0] || r.x > lim[1])
    r.ct = __builtin_inf();
  else if (r.y < lim[2] || r.y > lim[3])
    r.ct = __builtin_inf();
}

__device__ __forceinline__ void ellip_ap(R6& r, cdouble* ax) {
#pragma clang fp contract(off)
  const double u = r.x / ax[0];
  const double v = r.y / ax[1];
  if ((u * u + v * v) >= 1.0) r.ct = __builtin_inf();
}

__device__ __forceinline__ void apertures(R6& r, cDevElem& h, cdouble* pool) {
  if (h.flags & F_RAP) rect_ap(r, pool + h.rap);
  if (h.flags & F_EAP) ellip_ap(r, pool + h.eap);
}

__device__ __forceinline__ void entry_xform(R6& r, cDevElem& h, cdouble* pool) {
  if (h.flags & F_T1) add_t(r, pool + h.t1);
  if (h.flags & F_R1) mul_r(r, pool + h.r1);
}

__device__ __forceinline__ void exit_xform(R6& r, cDevElem& h, cdouble* pool) {
  if (h.flags & F_R2) mul_r(r, pool + h.r2);
  if (h.flags & F_T2) add_t(r, pool + h.t2);
}

#if !ATX_STRICT
// Fast build: 1/d from the hardware reciprocal estimate (v_rcp_f64) and two
// Newton steps q <- q + q (1 - d q), each of which squares the relative
// error of the estimate; the result is within an ulp of 1/d.
__device__ __forceinline__ double fast_rcp(double d) {
  double q = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, q, 1.0);
  q = __builtin_fma(q, e, q);
  e = __builtin_fma(-d, q, 1.0);
  return __builtin_fma(q, e, q);
}
#endif

// Full drift with its own 1/(1+delta) (atlalib.c:59-68, driftkickrad.c:36-45,
// diff_drift.c:30-43 share this arithmetic); the fast build takes fast_rcp
// for the division.
__device__ __forceinline__ void drift6(R6& r, double len) {
#if ATX_STRICT
  const double pn = 1.0 / (1.0 + r.dp);
#else
  const double pn = fast_rcp(1.0 + r.dp);
#endif
  const double nl = len * pn;
  r.x += nl * r.px;
  r.y += nl * r.py;
  r.ct += nl * pn * (r.px * r.px + r.py * r.py) / 2.0;
}

// In-magnet drift with nl = L/(1+delta) hoisted (driftkick.c:37-48).
// h2 = 1/(2(1+delta)) is used by the fast build only.
__device__ __forceinline__ void fastdrift(R6& r, double nl, double h2) {
#if ATX_STRICT
  (void)h2;
  r.x += nl * r.px;
  r.y += nl * r.py;
  r.ct += nl * (r.px * r.px + r.py * r.py) / (2.0 * (1.0 + r.dp));
#else
  r.x = __builtin_fma(nl, r.px, r.x);
  r.y = __builtin_fma(nl, r.py, r.y);
  const double p2 = __builtin_fma(r.px, r.px, r.py * r.py);
  r.ct = __builtin_fma(nl * p2, h2, r.ct);
#endif
}

// Addends of the first Horner step, B[order-1] and A[order-1] (0 for order
// 0), taken once per magnet. That step multiplies the constants B[order],
// A[order] by a coordinate and adds another constant, and an FMA reads at
// most one SGPR: the fast build therefore keeps these two addends in VGPRs
// for the whole magnet instead of moving them from SGPRs at every kick.
template <int N>
__device__ __forceinline__ void horner_addends(cdouble* a, cdouble* b, int n, double& blo,
                                               double& alo) {
  const int order = (N >= 0) ? N : n;
  blo = (order >= 1) ? b[order - 1] : 0.0;
  alo = (order >= 1) ? a[order - 1] : 0.0;
#if !ATX_STRICT
  if (N >= 1) {
    asm("" : "=v"(blo) : "0"(blo));
    asm("" : "=v"(alo) : "0"(alo));
  }
#endif
}

// Horner evaluation of sum (B_k + i A_k)(x + i y)^k (driftkick.c:92-110).
// N >= 0: order known at compile time; N < 0: runtime order n. blo, alo:
// B[order-1], A[order-1] from horner_addends (the same values).
template <int N>
__device__ __forceinline__ void field(const R6& r, cdouble* a, cdouble* b, int n, double blo,
                                      double alo, double& re, double& im) {
  const int order = (N >= 0) ? N : n;
  re = b[order];
  im = a[order];
#pragma unroll
  for (int k = order - 1; k >= 0; k--) {
    const double bk = (N >= 1 && k == order - 1) ? blo : b[k];
    const double ak = (N >= 1 && k == order - 1) ? alo : a[k];
#if ATX_STRICT
    const double t = re * r.x - im * r.y + bk;
    im = im * r.x + re * r.y + ak;
#else
    const double t = __builtin_fma(re, r.x, __builtin_fma(-im, r.y, bk));
    im = __builtin_fma(im, r.x, __builtin_fma(re, r.y, ak));
#endif
    re = t;
  }
}

// driftkick.c:92-115
template <int N>
__device__ __forceinline__ void str_kick(R6& r, cdouble* a, cdouble* b, int n, double blo,
                                         double alo, double kl) {
  double re, im;
  field<N>(r, a, b, n, blo, alo, re, im);
  r.px -= kl * re;
  r.py += kl * im;
}

// driftkick.c:51-89; kli = kl*irho prepared on the host
template <int N>
__device__ __forceinline__ void bnd_kick(R6& r, cdouble* a, cdouble* b, int n, double blo,
                                         double alo, double kl, double irho, double kli) {
  double re, im;
  field<N>(r, a, b, n, blo, alo, re, im);
#if ATX_STRICT
  r.px -= kl * (re - (r.dp - r.x * irho) * irho);
  r.py += kl * im;
  r.ct += kli * r.x;
#else
  const double u = __builtin_fma(-r.x, irho, r.dp);
  const double v = __builtin_fma(-u, irho, re);
  r.px = __builtin_fma(-kl, v, r.px);
  r.py = __builtin_fma(kl, im, r.py);
  r.ct = __builtin_fma(kli, r.x, r.ct);
#endif
}

// Radiating drift (diff_drift.c:30-43).
__device__ __forceinline__ void rad_drift(R6& r, double len) {
  const double pn = 1.0 / (1.0 + r.dp);
  const double nl = len * pn;
  r.x += nl * r.px;
  r.y += nl * r.py;
  r.ct += nl * pn * (r.px * r.px + r.py * r.py) / 2.0;
}

#if !ATX_STRICT
// One Newton step: for a reciprocal that only scales the radiated energy (a
// relative change of 1e-8 or less per kick), where an error of the squared
// estimate error is far below the rounding of delta.
__device__ __forceinline__ double rcp1(double d) {
  const double q = __builtin_amdgcn_rcp(d);
  return __builtin_fma(q, __builtin_fma(-d, q, 1.0), q);
}

// Fast build: reciprocal of 1 + delta after a kick from the one before it,
// pn = 1/(1 + delta_before). With the radiated fraction e = 1 - (1 + delta) pn
// (about 1e-7 per kick), 1/(1 + delta) = pn/(1 - e) = pn (1 + e + e^2) to a
// relative e^3, below 1e-17 for |e| <= 2e-6. A wave in which any lane
// radiated more than that in one kick takes fast_rcp.
__device__ __forceinline__ double rcp_after_kick(double opd2, double pn) {
  const double e = __builtin_fma(-opd2, pn, 1.0);
  double q = __builtin_fma(pn, __builtin_fma(e, e, e), pn);
  const bool far = __builtin_fabs(e) > 2e-6;
  if (__builtin_expect(__ballot(far) != 0, 0)) {
    if (far) q = fast_rcp(opd2);
  }
  return q;
}

// Fast build: inside a radiating magnet delta changes only in the kicks, so
// pn = 1/(1+delta) is computed once at entry and then by each kick for the
// radiated delta, and carried to the drifts.
__device__ __forceinline__ void rad_drift_pn(R6& r, double len, double pn) {
  const double nl = len * pn;
  r.x = __builtin_fma(nl, r.px, r.x);
  r.y = __builtin_fma(nl, r.py, r.y);
  const double p2 = __builtin_fma(r.px, r.px, r.py * r.py);
  r.ct = __builtin_fma(nl * pn * 0.5, p2, r.ct);
}
#endif

// Radiating straight kick (diff_str_kick.c:19-83, bdiff == NULL).
// radc = RAD_CONST*gamma^3 (strict); radk = radc*kl (fast). The fast build
// takes pn = 1/(1+delta) at entry and returns it updated to the radiated
// delta.
template <int N>
__device__ __forceinline__ void rad_str_kick(R6& r, cdouble* a, cdouble* b, int n, double blo,
                                             double alo, double kl, double radc, double radk,
                                             double& pn_io) {
  double re, im;
  field<N>(r, a, b, n, blo, alo, re, im);
#if ATX_STRICT
  (void)pn_io;
  (void)radk;
  double pn = 1.0 / (1.0 + r.dp);
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  const double vn2 = 1.0 / (1.0 + xpr * xpr + ypr * ypr);
  const double cr = im * ypr - re * xpr;
  const double b2p = (re * re + im * im + cr * cr) * vn2;
  const double factor = (1.0 + (xpr * xpr + ypr * ypr) / 2.0) / (pn * pn) * kl;
  r.dp -= radc * b2p * factor;
  pn = 1.0 / (1.0 + r.dp);
  r.px = xpr / pn;
  r.py = ypr / pn;
  r.px -= kl * re;
  r.py += kl * im;
#else
  (void)radc;
  const double pn_in = pn_io;
  const double opd = 1.0 + r.dp;
  const double xpr = r.px * pn_in;
  const double ypr = r.py * pn_in;
  const double s2 = __builtin_fma(xpr, xpr, ypr * ypr);
  const double cr = __builtin_fma(im, ypr, -re * xpr);
  const double b2 = __builtin_fma(cr, cr, __builtin_fma(re, re, im * im));
  // g = (1+s2/2)/(1+s2) = 1 - (s2/2)/(1+s2), with 1 - s2 + s2^2 for the
  // reciprocal: relative error s2^4/2 in g, below 1e-16 for s2 <= 1e-4
  // (angles up to 10 mrad); a wave with a wider angle divides.
  double g = __builtin_fma(-0.5 * s2, __builtin_fma(s2, s2, 1.0 - s2), 1.0);
  const bool wide = s2 > 1e-4;
  if (__builtin_expect(__ballot(wide) != 0, 0)) {
    if (wide) g = __builtin_fma(0.5, s2, 1.0) * rcp1(1.0 + s2);
  }
  r.dp = __builtin_fma(-radk, (b2 * g) * (opd * opd), r.dp);
  const double opd2 = 1.0 + r.dp;
  r.px = __builtin_fma(xpr, opd2, -kl * re);
  r.py = __builtin_fma(ypr, opd2, kl * im);
  pn_io = rcp_after_kick(opd2, pn_in);
#endif
}

// Radiating curved-frame kick (diff_bnd_kick.c:85-174, bdiff == NULL).
template <int N>
__device__ __forceinline__ void rad_bnd_kick(R6& r, cdouble* a, cdouble* b, int n, double blo,
                                             double alo, double kl, double irho, double kli,
                                             double radc, double radk, double& pn_io) {
  double re, im;
  field<N>(r, a, b, n, blo, alo, re, im);
#if ATX_STRICT
  (void)pn_io;
  (void)radk;
  double pn = 1.0 / (1.0 + r.dp);
  const double x = r.x;
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  const double by = re + irho;
  const double bx = im;
  const double hx = 1.0 + x * irho;
  const double vn2 = 1.0 / (hx * hx + xpr * xpr + ypr * ypr);
  const double byh = by * hx;
  const double bxh = bx * hx;
  const double cr = bx * ypr - by * xpr;
  const double b2p = (byh * byh + bxh * bxh + cr * cr) * vn2;
  const double factor = (1.0 + x * irho + (xpr * xpr + ypr * ypr) / 2.0) / (pn * pn) * kl;
  const double d0 = r.dp;
  r.dp -= radc * b2p * factor;
  pn = 1.0 / (1.0 + r.dp);
  r.px = xpr / pn;
  r.py = ypr / pn;
  r.px -= kl * (re - (d0 - r.x * irho) * irho);
  r.py += kl * im;
  r.ct += kli * r.x;
#else
  (void)radc;
  const double pn_in = pn_io;
  const double x = r.x;
  const double opd = 1.0 + r.dp;
  const double xpr = r.px * pn_in;
  const double ypr = r.py * pn_in;
  const double by = re + irho;
  const double hx = __builtin_fma(x, irho, 1.0);
  const double s2 = __builtin_fma(xpr, xpr, ypr * ypr);
  const double byh = by * hx;
  const double bxh = im * hx;
  const double cr = __builtin_fma(im, ypr, -by * xpr);
  const double b2 = __builtin_fma(cr, cr, __builtin_fma(byh, byh, bxh * bxh));
  const double g = __builtin_fma(0.5, s2, hx) * rcp1(__builtin_fma(hx, hx, s2));
  const double d0 = r.dp;
  r.dp = __builtin_fma(-radk, (b2 * g) * (opd * opd), r.dp);
  const double opd2 = 1.0 + r.dp;
  const double u = __builtin_fma(-x, irho, d0);
  const double v = __builtin_fma(-u, irho, re);
  r.px = __builtin_fma(xpr, opd2, -kl * v);
  r.py = __builtin_fma(ypr, opd2, kl * im);
  r.ct = __builtin_fma(kli, x, r.ct);
  pn_io = rcp_after_kick(opd2, pn_in);
#endif
}

// Edge focusing (atphyslib.c:15-79, diff_bend_fringe.c:2-53). sgn = +1 at
// the entrance, -1 at the exit; fx, c0 = irho*tan(edge), edge and corr/psi
// come from the scalar block.
__device__ __noinline__ double edge_fy_tan(int mode, double dp, double px, double irho,
                                           double edge_angle, double corr, double sgn) {
  switch (mode) {
    case E_TAN1:
      return irho * tan(edge_angle - corr / (1.0 + dp));
    case E_TAN2:
      return irho * tan(edge_angle - corr / (1.0 + dp)) / (1.0 + dp);
    case E_TAN3:
      return (sgn > 0.0) ? irho * tan(edge_angle - corr + px / (1.0 + dp))
                         : irho * tan(edge_angle - corr - px / (1.0 + dp));
    case E_RTAN1: {
      const double pn = 1.0 / (1.0 + dp);
      return irho * tan(edge_angle - corr * pn);
    }
    case E_RTAN2: {
      const double pn = 1.0 / (1.0 + dp);
      return irho * tan(edge_angle - corr * pn) * pn;
    }
    default: {  // E_RTAN3
      const double pn = 1.0 / (1.0 + dp);
      return irho * tan(edge_angle - corr + sgn * px * pn);
    }
  }
}

__device__ __forceinline__ void edge(R6& r, int mode, double irho, double fx, double c0,
                                     double edge_angle, double corr, double sgn) {
  double fy;
  if (mode == E_CONST)
    fy = c0;
  else if (mode == E_DIVP)
    fy = c0 / (1.0 + r.dp);
  else if (mode == E_MULPN)
    fy = c0 * (1.0 / (1.0 + r.dp));
  else
    fy = edge_fy_tan(mode, r.dp, r.px, irho, edge_angle, corr, sgn);
  r.px += r.x * fx;
  r.py -= r.y * fy;
}

// Lee-Whiting hard-edge quadrupole fringe (quadfringe.c:3-51).
__device__ __forceinline__ void quad_fringe_inl(R6& r, double b2, bool entrance) {
  const double u = b2 / (12.0 * (1.0 + r.dp));
  const double x2 = r.x * r.x;
  const double z2 = r.y * r.y;
  const double xz = r.x * r.y;
  const double gx = u * (x2 + 3.0 * z2) * r.x;
  const double gz = u * (z2 + 3.0 * x2) * r.y;
  const double dpx = 3.0 * u * (2.0 * xz * r.py - (x2 + z2) * r.px);
  const double dpy = 3.0 * u * (2.0 * xz * r.px - (x2 + z2) * r.py);
  if (entrance) {
    r.x += gx;
    r.y -= gz;
    r.ct -= (gz * r.py - gx * r.px) / (1.0 + r.dp);
    r.px += dpx;
    r.py -= dpy;
  } else {
    r.x -= gx;
    r.y += gz;
    r.ct += (gz * r.py - gx * r.px) / (1.0 + r.dp);
    r.px -= dpx;
    r.py += dpy;
  }
}

// 2x2 linear blocks of the elegant fringe (quadfringe.c:93-163), matrices
// prepared on the host: m = {Rxx, Rxp, Rpx, Rpp, Ryy, Ryq, Rqy, Rqq}.
__device__ __forceinline__ void apply_2x2(R6& r, cdouble* m) {
  const double x = r.x, px = r.px, y = r.y, py = r.py;
  r.x = m[0] * x + m[1] * px;
  r.px = m[2] * x + m[3] * px;
  r.y = m[4] * y + m[5] * py;
  r.py = m[6] * y + m[7] * py;
}

// mode: Q_LEE or Q_ELEGANT; m: the two 2x2 blocks around the Lee-Whiting map
__device__ __noinline__ R6 quad_fringe_v(R6 r, int mode, cdouble* m, double b2, bool entrance) {
  if (mode == Q_ELEGANT) apply_2x2(r, m);
  quad_fringe_inl(r, b2, entrance);
  if (mode == Q_ELEGANT) apply_2x2(r, m + 8);
  return r;
}

__device__ __forceinline__ void quad_fringe_in(R6& r, cDevElem& h, cdouble* pool, double b2) {
  r = quad_fringe_v(r, h.fringe & 0xff, pool + h.fq, b2, true);
}

__device__ __forceinline__ void quad_fringe_out(R6& r, cDevElem& h, cdouble* pool, double b2) {
  r = quad_fringe_v(r, (h.fringe >> 8) & 0xff, pool + h.fq + 16, b2, false);
}

// Slices of a non-default integrator (atgpu/SymplecticIntegrator.cpp types
// 1-3, 5, 6): the host-built op table lists the drifts and kicks of one
// slice; kept out of line, generic in the multipole order.
template <bool BEND, bool RAD>
__device__ __noinline__ R6 slices_ops(R6 r, cdouble* a, cdouble* b, int n, cdouble* op, int nslice,
                                      double pn0, double irho, double radc) {
  const int nop = (int)op[0];
  double blo, alo;
  horner_addends<-1>(a, b, n, blo, alo);
#if ATX_STRICT
  const double h2 = 0.0;
#else
  const double h2 = 0.5 / (1.0 + r.dp);
#endif
  for (int m = 0; m < nslice; m++) {
    for (int j = 0; j < nop; j++) {
      const double len = op[2 + 2 * j];
      if (op[1 + 2 * j] == 0.0) {
        if (RAD) rad_drift(r, len); else fastdrift(r, len * pn0, h2);
      } else {
        if (RAD) {
          double pn = 1.0 / (1.0 + r.dp);
          if (BEND) rad_bnd_kick<-1>(r, a, b, n, blo, alo, len, irho, len * irho, radc, radc * len, pn);
          else rad_str_kick<-1>(r, a, b, n, blo, alo, len, radc, radc * len, pn);
        } else {
          if (BEND) bnd_kick<-1>(r, a, b, n, blo, alo, len, irho, len * irho); else str_kick<-1>(r, a, b, n, blo, alo, len);
        }
      }
    }
  }
  return r;
}

// Particles per lane: the element code below takes the P particles a lane
// tracks as an array and runs every step of a magnet for all of them before
// the next step, so the P independent dependency chains interleave in one
// instruction stream, and the magnet's scalar constants, loop control and
// branches serve P particles. Per-particle maps that are rare or large stay
// out of line and run once per particle.
#if !ATX_STRICT
// In-magnet drift of the fast build: nlh = nl / (2 (1+delta)) hoisted with
// nl (delta is constant in a non-radiating magnet).
__device__ __forceinline__ void fastdrift_h(R6& r, double nl, double nlh) {
  r.x = __builtin_fma(nl, r.px, r.x);
  r.y = __builtin_fma(nl, r.py, r.y);
  const double p2 = __builtin_fma(r.px, r.px, r.py * r.py);
  r.ct = __builtin_fma(nlh, p2, r.ct);
}

// Radiating magnets of the fast build carry the transverse momenta as
// angles x' = px/(1+delta), y' = py/(1+delta) from entry to exit: a drift is
// then x += L x', ct += (L/2)(x'^2 + y'^2), its x'^2 + y'^2 is the kick's
// s2 (the kick that follows a drift sees the same angles), and a kick turns
// px = x'(1+delta') - kl Re, the reference's update, into
// x' <- x' - kl Re / (1+delta'). pn = 1/(1+delta) as in rad_str_kick.
struct Ang {
  double xp, yp, s2, pn;
};

__device__ __forceinline__ Ang to_angles(const R6& r) {
  const double pn = fast_rcp(1.0 + r.dp);
  return Ang{r.px * pn, r.py * pn, 0.0, pn};
}

__device__ __forceinline__ void from_angles(R6& r, const Ang& g) {
  const double opd = 1.0 + r.dp;
  r.px = g.xp * opd;
  r.py = g.yp * opd;
}

__device__ __forceinline__ void rad_drift_ang(R6& r, Ang& g, double len, double hlen) {
  r.x = __builtin_fma(len, g.xp, r.x);
  r.y = __builtin_fma(len, g.yp, r.y);
  g.s2 = __builtin_fma(g.xp, g.xp, g.yp * g.yp);
  r.ct = __builtin_fma(hlen, g.s2, r.ct);
}

// diff_str_kick.c:19-83 (bdiff == NULL) in angle form; radk = RAD_CONST
// gamma^3 kl. The radiated factor (1+s2/2)/(1+s2) is the series of
// rad_str_kick (relative error s2^4/2; a wave with s2 > 1e-4 divides).
template <int N>
__device__ __forceinline__ void rad_str_kick_ang(R6& r, Ang& g, cdouble* a, cdouble* b, int n,
                                                 double blo, double alo, double kl, double radk) {
  double re, im;
  field<N>(r, a, b, n, blo, alo, re, im);
  const double s2 = g.s2;
  const double opd = 1.0 + r.dp;
  const double cr = __builtin_fma(im, g.yp, -re * g.xp);
  const double b2 = __builtin_fma(cr, cr, __builtin_fma(re, re, im * im));
  // 1 - (s2/2)(1 - s2 + s2^2) by Horner in s2
  const double f = __builtin_fma(s2, __builtin_fma(s2, __builtin_fma(-0.5, s2, 0.5), -0.5), 1.0);
  const double oo = opd * opd;
  const double d0 = r.dp;
  double dp = __builtin_fma(-radk, (b2 * f) * oo, d0);
  double opd2 = 1.0 + dp;
  double e = __builtin_fma(-opd2, g.pn, 1.0);
  double pn = __builtin_fma(g.pn, __builtin_fma(e, e, e), g.pn);
  // one vote for both exceptions of the wave: a wide angle (the factor by
  // division) and a large radiated fraction (the reciprocal from fast_rcp)
  const bool wide = s2 > 1e-4;
  if (__builtin_expect(__ballot(wide || __builtin_fabs(e) > 2e-6) != 0, 0)) {
    if (wide) {
      const double fw = __builtin_fma(0.5, s2, 1.0) * rcp1(1.0 + s2);
      dp = __builtin_fma(-radk, (b2 * fw) * oo, d0);
      opd2 = 1.0 + dp;
      e = __builtin_fma(-opd2, g.pn, 1.0);
      pn = __builtin_fma(g.pn, __builtin_fma(e, e, e), g.pn);
    }
    if (__builtin_fabs(e) > 2e-6) pn = fast_rcp(opd2);
  }
  r.dp = dp;
  const double klpn = kl * pn;
  g.xp = __builtin_fma(-re, klpn, g.xp);
  g.yp = __builtin_fma(im, klpn, g.yp);
  g.pn = pn;
}

// diff_bnd_kick.c:85-174 (bdiff == NULL) in angle form.
template <int N>
__device__ __forceinline__ void rad_bnd_kick_ang(R6& r, Ang& g, cdouble* a, cdouble* b, int n,
                                                 double blo, double alo, double kl, double irho,
                                                 double kli, double radk) {
  double re, im;
  field<N>(r, a, b, n, blo, alo, re, im);
  const double x = r.x;
  const double s2 = g.s2;
  const double d0 = r.dp;
  const double opd = 1.0 + d0;
  const double by = re + irho;
  const double hx = __builtin_fma(x, irho, 1.0);
  const double byh = by * hx;
  const double bxh = im * hx;
  const double cr = __builtin_fma(im, g.yp, -by * g.xp);
  const double b2 = __builtin_fma(cr, cr, __builtin_fma(byh, byh, bxh * bxh));
  const double f = __builtin_fma(0.5, s2, hx) * rcp1(__builtin_fma(hx, hx, s2));
  r.dp = __builtin_fma(-radk, (b2 * f) * (opd * opd), d0);
  const double pn = rcp_after_kick(1.0 + r.dp, g.pn);
  const double u = __builtin_fma(-x, irho, d0);
  const double v = __builtin_fma(-u, irho, re);
  const double klpn = kl * pn;
  g.xp = __builtin_fma(-v, klpn, g.xp);
  g.yp = __builtin_fma(im, klpn, g.yp);
  r.ct = __builtin_fma(kli, x, r.ct);
  g.pn = pn;
}
#endif

// Entry and exit of a thick multipole, per particle: FieldScaling, T1/R1,
// apertures, bend edge, quadrupole fringe (StrMPoleSymplectic4Pass.c:57-78,
// 96-104; BndMPoleSymplectic4Pass.c:75-96, 114-122).
// One test for the transforms and apertures an unmisaligned magnet without
// apertures has none of, before the tests of each.
constexpr uint32_t F_ENTRY_ANY = F_T1 | F_R1 | F_RAP | F_EAP;
constexpr uint32_t F_EXIT_ANY = F_R2 | F_T2 | F_RAP | F_EAP | F_SCALE;

template <bool BEND>
__device__ __forceinline__ void mpole_entry(R6& r, cDevElem& h, cdouble* pool, cdouble* s) {
  if (h.flags & F_ENTRY_ANY) {
    entry_xform(r, h, pool);
    apertures(r, h, pool);
  }
  if (BEND) edge(r, h.edge & 0xff, s[S_IRHO], s[S_FX_IN], s[S_FY_IN], s[S_EDGE_IN], s[S_CORR_IN], 1.0);
  if (h.fringe & 0xff) quad_fringe_in(r, h, pool, s[S_B2]);
}

template <bool BEND>
__device__ __forceinline__ void mpole_exit(R6& r, cDevElem& h, cdouble* pool, cdouble* s) {
  if (h.fringe & 0xff00) quad_fringe_out(r, h, pool, s[S_B2]);
  if (BEND) edge(r, (h.edge >> 8) & 0xff, s[S_IRHO], s[S_FX_OUT], s[S_FY_OUT], s[S_EDGE_OUT], s[S_CORR_OUT], -1.0);
  if (h.flags & F_EXIT_ANY) {
    apertures(r, h, pool);
    exit_xform(r, h, pool);
    if (h.flags & F_SCALE) change_pref(r, s[S_INVSCALE]);
  }
}

// StrMPoleSymplectic4Pass.c:57-104 and BndMPoleSymplectic4Pass.c:75-122 for
// the P particles of a lane; U slices per loop iteration.
template <bool BEND, int N, int P, int U>
__device__ __forceinline__ void mpole(R6* r, cDevElem& h, cdouble* pool) {
  cdouble* s = pool + h.sc;
  cdouble* a = pool + h.poly;
  cdouble* b = a + h.order + 1;
  const int n = h.order;
  double pn[P], nl1[P], nl2[P];
#pragma unroll
  for (int q = 0; q < P; q++) {
    if (h.flags & F_SCALE) change_pref(r[q], s[S_SCALE]);
#if ATX_STRICT
    pn[q] = 1.0 / (1.0 + r[q].dp);  // before T1/R1, as the reference hoists it
#else
    pn[q] = fast_rcp(1.0 + r[q].dp);  // before T1/R1, as the reference hoists it
#endif
    nl1[q] = s[S_L1] * pn[q];
    nl2[q] = s[S_L2] * pn[q];
    mpole_entry<BEND>(r[q], h, pool, s);
  }
  const double irho = s[S_IRHO];
  if (h.ops >= 0) {
#pragma unroll
    for (int q = 0; q < P; q++)
      r[q] = slices_ops<BEND, false>(r[q], a, b, n, pool + h.ops, h.nslice, pn[q], irho, 0.0);
  } else {
    const double k1 = s[S_K1], k2 = s[S_K2];
    const double ki1 = k1 * irho, ki2 = k2 * irho;
    double blo, alo;
    horner_addends<N>(a, b, n, blo, alo);
#if ATX_STRICT
    // U slices per iteration: a constant-count inner loop, fully unrolled
    // (the slices may hold wave-wide votes, which a runtime-count unroll
    // cannot duplicate), then the remaining slices one by one
    auto slice = [&]() {
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift(r[q], nl1[q], 0.0);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) bnd_kick<N>(r[q], a, b, n, blo, alo, k1, irho, ki1); else str_kick<N>(r[q], a, b, n, blo, alo, k1);
      }
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift(r[q], nl2[q], 0.0);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) bnd_kick<N>(r[q], a, b, n, blo, alo, k2, irho, ki2); else str_kick<N>(r[q], a, b, n, blo, alo, k2);
      }
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift(r[q], nl2[q], 0.0);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) bnd_kick<N>(r[q], a, b, n, blo, alo, k1, irho, ki1); else str_kick<N>(r[q], a, b, n, blo, alo, k1);
      }
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift(r[q], nl1[q], 0.0);
    };
    int m = 0;
    for (; m + U <= h.nslice; m += U) {
#pragma unroll
      for (int u = 0; u < U; u++) slice();
    }
    for (; m < h.nslice; m++) slice();
#else
    double nh1[P], nh2[P];  // delta is constant through the slices
#pragma unroll
    for (int q = 0; q < P; q++) {
      const double h2 = 0.5 * fast_rcp(1.0 + r[q].dp);
      nh1[q] = nl1[q] * h2;
      nh2[q] = nl2[q] * h2;
    }
    // U slices per iteration: a constant-count inner loop, fully unrolled
    // (the slices may hold wave-wide votes, which a runtime-count unroll
    // cannot duplicate), then the remaining slices one by one
    auto slice = [&]() {
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift_h(r[q], nl1[q], nh1[q]);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) bnd_kick<N>(r[q], a, b, n, blo, alo, k1, irho, ki1); else str_kick<N>(r[q], a, b, n, blo, alo, k1);
      }
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift_h(r[q], nl2[q], nh2[q]);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) bnd_kick<N>(r[q], a, b, n, blo, alo, k2, irho, ki2); else str_kick<N>(r[q], a, b, n, blo, alo, k2);
      }
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift_h(r[q], nl2[q], nh2[q]);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) bnd_kick<N>(r[q], a, b, n, blo, alo, k1, irho, ki1); else str_kick<N>(r[q], a, b, n, blo, alo, k1);
      }
#pragma unroll
      for (int q = 0; q < P; q++) fastdrift_h(r[q], nl1[q], nh1[q]);
    };
    int m = 0;
    for (; m + U <= h.nslice; m += U) {
#pragma unroll
      for (int u = 0; u < U; u++) slice();
    }
    for (; m < h.nslice; m++) slice();
#endif
  }
#pragma unroll
  for (int q = 0; q < P; q++) mpole_exit<BEND>(r[q], h, pool, s);
}

// StrMPoleSymplectic4RadPass.c:61-110 and BndMPoleSymplectic4RadPass.c:76-123
// for the P particles of a lane; U slices per loop iteration.
template <bool BEND, int N, int P, int U>
__device__ __forceinline__ void mpole_rad(R6* r, cDevElem& h, cdouble* pool) {
  cdouble* s = pool + h.sc;
  cdouble* a = pool + h.poly;
  cdouble* b = a + h.order + 1;
  const int n = h.order;
#pragma unroll
  for (int q = 0; q < P; q++) {
    if (h.flags & F_SCALE) change_pref(r[q], s[S_SCALE]);
    mpole_entry<BEND>(r[q], h, pool, s);
  }
  const double irho = s[S_IRHO];
  const double radc = s[S_RADC];
  if (h.ops >= 0) {
#pragma unroll
    for (int q = 0; q < P; q++)
      r[q] = slices_ops<BEND, true>(r[q], a, b, n, pool + h.ops, h.nslice, 0.0, irho, radc);
  } else {
    const double l1 = s[S_L1], l2 = s[S_L2], k1 = s[S_K1], k2 = s[S_K2];
    const double ki1 = k1 * irho, ki2 = k2 * irho;
    double blo, alo;
    horner_addends<N>(a, b, n, blo, alo);
#if ATX_STRICT
    double pn = 0.0;  // unused: the strict steps divide as the reference does
    // U slices per iteration: a constant-count inner loop, fully unrolled
    // (the slices may hold wave-wide votes, which a runtime-count unroll
    // cannot duplicate), then the remaining slices one by one
    auto slice = [&]() {
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift(r[q], l1);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) rad_bnd_kick<N>(r[q], a, b, n, blo, alo, k1, irho, ki1, radc, 0.0, pn);
        else rad_str_kick<N>(r[q], a, b, n, blo, alo, k1, radc, 0.0, pn);
      }
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift(r[q], l2);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) rad_bnd_kick<N>(r[q], a, b, n, blo, alo, k2, irho, ki2, radc, 0.0, pn);
        else rad_str_kick<N>(r[q], a, b, n, blo, alo, k2, radc, 0.0, pn);
      }
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift(r[q], l2);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) rad_bnd_kick<N>(r[q], a, b, n, blo, alo, k1, irho, ki1, radc, 0.0, pn);
        else rad_str_kick<N>(r[q], a, b, n, blo, alo, k1, radc, 0.0, pn);
      }
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift(r[q], l1);
    };
    int m = 0;
    for (; m + U <= h.nslice; m += U) {
#pragma unroll
      for (int u = 0; u < U; u++) slice();
    }
    for (; m < h.nslice; m++) slice();
#else
    const double rk1 = s[S_RK1], rk2 = s[S_RK2], hl1 = s[S_HL1], hl2 = s[S_HL2];
    Ang g[P];
#pragma unroll
    for (int q = 0; q < P; q++) g[q] = to_angles(r[q]);  // one reciprocal per magnet
    // U slices per iteration: a constant-count inner loop, fully unrolled
    // (the slices may hold wave-wide votes, which a runtime-count unroll
    // cannot duplicate), then the remaining slices one by one
    auto slice = [&]() {
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift_ang(r[q], g[q], l1, hl1);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) rad_bnd_kick_ang<N>(r[q], g[q], a, b, n, blo, alo, k1, irho, ki1, rk1);
        else rad_str_kick_ang<N>(r[q], g[q], a, b, n, blo, alo, k1, rk1);
      }
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift_ang(r[q], g[q], l2, hl2);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) rad_bnd_kick_ang<N>(r[q], g[q], a, b, n, blo, alo, k2, irho, ki2, rk2);
        else rad_str_kick_ang<N>(r[q], g[q], a, b, n, blo, alo, k2, rk2);
      }
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift_ang(r[q], g[q], l2, hl2);
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (BEND) rad_bnd_kick_ang<N>(r[q], g[q], a, b, n, blo, alo, k1, irho, ki1, rk1);
        else rad_str_kick_ang<N>(r[q], g[q], a, b, n, blo, alo, k1, rk1);
      }
#pragma unroll
      for (int q = 0; q < P; q++) rad_drift_ang(r[q], g[q], l1, hl1);
    };
    int m = 0;
    for (; m + U <= h.nslice; m += U) {
#pragma unroll
      for (int u = 0; u < U; u++) slice();
    }
    for (; m < h.nslice; m++) slice();
#pragma unroll
    for (int q = 0; q < P; q++) from_angles(r[q], g[q]);
#endif
  }
#pragma unroll
  for (int q = 0; q < P; q++) mpole_exit<BEND>(r[q], h, pool, s);
}

template <bool BEND, bool RAD, int P, int U>
__device__ __forceinline__ void mpole_dispatch(R6* r, cDevElem& h, cdouble* pool) {
  switch (h.order) {
    case 0: RAD ? mpole_rad<BEND, 0, P, U>(r, h, pool) : mpole<BEND, 0, P, U>(r, h, pool); break;
    case 1: RAD ? mpole_rad<BEND, 1, P, U>(r, h, pool) : mpole<BEND, 1, P, U>(r, h, pool); break;
    case 2: RAD ? mpole_rad<BEND, 2, P, U>(r, h, pool) : mpole<BEND, 2, P, U>(r, h, pool); break;
    case 3: RAD ? mpole_rad<BEND, 3, P, U>(r, h, pool) : mpole<BEND, 3, P, U>(r, h, pool); break;
    default: RAD ? mpole_rad<BEND, -1, P, U>(r, h, pool) : mpole<BEND, -1, P, U>(r, h, pool); break;
  }
}

// ThinMPolePass.c:42-63
__device__ __forceinline__ void thin(R6& r, cDevElem& h, cdouble* pool) {
  cdouble* s = pool + h.sc;
  cdouble* a = pool + h.poly;
  cdouble* b = a + h.order + 1;
  if (h.flags & F_SCALE) change_pref(r, s[S_SCALE]);
  entry_xform(r, h, pool);
  apertures(r, h, pool);
  switch (h.order) {
    case 0: str_kick<0>(r, a, b, 0, 0.0, 0.0, 1.0); break;
    case 1: str_kick<1>(r, a, b, 1, b[0], a[0], 1.0); break;
    case 2: str_kick<2>(r, a, b, 2, b[1], a[1], 1.0); break;
    case 3: str_kick<3>(r, a, b, 3, b[2], a[2], 1.0); break;
    default: str_kick<-1>(r, a, b, h.order, 0.0, 0.0, 1.0); break;
  }
  const double bax = s[S_BAX], bay = s[S_BAY];
  r.px += bax * r.dp;
  r.py -= bay * r.dp;
  r.ct -= bax * r.x - bay * r.y;
  exit_xform(r, h, pool);
  if (h.flags & F_SCALE) change_pref(r, s[S_INVSCALE]);
}

// attrackfunc.c:5-42: hft = h/f - T0 prepared on the host. The sine is the
// host C library's, restated bit for bit (atx_libm.h), in both builds.
__device__ __noinline__ double rf_kick(double ct, double nv, double twopif, double lag,
                                       double hft0, double philag, double nt) {
#pragma clang fp contract(off)
  return -nv * libm::host_sin(twopif * ((ct - lag) / 2.99792458e8 - hft0 * nt) - philag);
}

__device__ __forceinline__ void rf_cavity(R6& r, cDevElem& h, cdouble* pool, int nturn) {
  cdouble* s = pool + h.sc;
  const double nt = (double)nturn;
  if (!(h.flags & F_THICK)) {
    if (h.flags & F_RFKICK)
      r.dp += rf_kick(r.ct, s[S_RF_NV], s[S_RF_TWOPIF], s[S_RF_LAG], s[S_RF_HFT0], s[S_RF_PHILAG], nt);
  } else {
    drift6(r, s[S_RF_HALF]);
    if (h.flags & F_RFKICK)
      r.dp += rf_kick(r.ct, s[S_RF_NV], s[S_RF_TWOPIF], s[S_RF_LAG], s[S_RF_HFT0], s[S_RF_PHILAG], nt);
    drift6(r, s[S_RF_HALF]);
  }
}

// CavityPass.c:18-56: no turn dependence (nv = V/E prepared on the host).
__device__ __noinline__ R6 cavity(R6 r, cdouble* s, bool thick) {
  if (thick) drift6(r, s[S_RF_HALF]);
  r.dp += -s[S_RF_NV] *
          libm::host_sin(s[S_RF_TWOPIF] * (r.ct - s[S_RF_LAG]) / 2.99792458e8 - s[S_RF_PHILAG]);
  if (thick) drift6(r, s[S_RF_HALF]);
  return r;
}

// Exact drift, absolute path length (exactdrift.c:4-16).
__device__ __forceinline__ void exact_drift(R6& r, double L) {
  const double opd = 1.0 + r.dp;
  const double pz = sqrt(opd * opd - r.px * r.px - r.py * r.py);
  const double nl = L / pz;
  r.x += r.px * nl;
  r.y += r.py * nl;
  r.ct += nl * opd;
}

// The same drift in the form of the radiating pass (diff_exactdrift.c:30-43).
__device__ __forceinline__ void exact_rad_drift(R6& r, double L) {
  const double pn = 1.0 / (1.0 + r.dp);
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  const double nl = L * pn / sqrt(1.0 - xpr * xpr - ypr * ypr);
  r.x += nl * r.px;
  r.y += nl * r.py;
  r.ct += nl * (1.0 + r.dp);
}

// Radiating straight kick with the exact velocity (diff_str_exactkick.c:14-71).
__device__ __forceinline__ void exact_rad_kick(R6& r, cdouble* a, cdouble* b, int n, double kl,
                                               double radc) {
  double re, im;
  field<-1>(r, a, b, n, 0.0, 0.0, re, im);
  double pn = 1.0 / (1.0 + r.dp);
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  const double t = im * xpr + re * ypr;
  const double b2p = im * im + re * re - t * t;
  const double factor = kl / (pn * pn) / sqrt(1.0 - xpr * xpr - ypr * ypr);
  r.dp -= radc * b2p * factor;
  pn = 1.0 / (1.0 + r.dp);
  r.px = xpr / pn;
  r.py = ypr / pn;
  r.px -= kl * re;
  r.py += kl * im;
}

// PTC multipole fringe (Forest 13.29) of every order, edge = +1 at the
// entrance, -1 at the exit (exactmultipolefringe.c:2-95, skip_b0 = 0).
__device__ __noinline__ R6 multipole_fringe(R6 r, cdouble* pa, cdouble* pb, int order, double edge) {
  double fx = 0, fy = 0, fx_x = 0, fx_y = 0, fy_x = 0, fy_y = 0;
  double rx = 1.0, ix = 0.0;
  for (int k = 0; k <= order; k++) {
    const double bn = pb[k], an = pa[k];
    const int j = k + 1;
    const double drx = rx, dix = ix;
    rx = drx * r.x - dix * r.y;
    ix = drx * r.y + dix * r.x;
    double u = bn * rx - an * ix;
    double v = bn * ix + an * rx;
    double du = bn * drx - an * dix;
    double dv = bn * dix + an * drx;
    const double f1 = -edge / 4.0 / (j + 1);
    u = u * f1;
    v = v * f1;
    du = du * f1;
    dv = dv * f1;
    const double dux = j * du, dvx = j * dv, duy = -j * dv, dvy = j * du;
    const double nf = 1.0 * (j + 2) / j;
    fx += u * r.x + nf * v * r.y;
    fy += u * r.y - nf * v * r.x;
    fx_x += dux * r.x + u + nf * r.y * dvx;
    fx_y += duy * r.x + nf * v + nf * r.y * dvy;
    fy_x += dux * r.y - nf * v - nf * r.x * dvx;
    fy_y += duy * r.y + u - nf * r.x * dvy;
  }
  const double del = 1.0 / (1 + r.dp);
  const double ma = 1 - fx_x * del;
  const double mb = -fy_x * del;
  const double md = 1 - fy_y * del;
  const double mc = -fx_y * del;
  r.x = r.x - fx * del;
  r.y = r.y - fy * del;
  const double pxf = (md * r.px - mb * r.py) / (ma * md - mb * mc);
  const double pyf = (ma * r.py - mc * r.px) / (ma * md - mb * mc);
  r.py = pyf;
  r.px = pxf;
  r.ct = r.ct - (r.px * fx + r.py * fy) * del * del;
  return r;
}

// ExactMultipolePass.c:27-96 / ExactMultipoleRadPass.c:26-104 slices, out
// of line (generic order): these passes are rare in rings.
__device__ __noinline__ R6 exact_slices(R6 r, cdouble* a, cdouble* b, int n, int nslice,
                                        cdouble* s, bool rad) {
  const double l1 = s[S_L1], l2 = s[S_L2], k1 = s[S_K1], k2 = s[S_K2];
  if (rad) {
    const double radc = s[S_RADC];
    for (int m = 0; m < nslice; m++) {
      exact_rad_drift(r, l1);
      exact_rad_kick(r, a, b, n, k1, radc);
      exact_rad_drift(r, l2);
      exact_rad_kick(r, a, b, n, k2, radc);
      exact_rad_drift(r, l2);
      exact_rad_kick(r, a, b, n, k1, radc);
      exact_rad_drift(r, l1);
    }
  } else {
    for (int m = 0; m < nslice; m++) {
      exact_drift(r, l1);
      str_kick<-1>(r, a, b, n, 0.0, 0.0, k1);
      exact_drift(r, l2);
      str_kick<-1>(r, a, b, n, 0.0, 0.0, k2);
      exact_drift(r, l2);
      str_kick<-1>(r, a, b, n, 0.0, 0.0, k1);
      exact_drift(r, l1);
    }
  }
  return r;
}

__device__ __forceinline__ void exact_mpole(R6& r, cDevElem& h, cdouble* pool, bool rad) {
  cdouble* s = pool + h.sc;
  cdouble* a = pool + h.poly;
  cdouble* b = a + h.order + 1;
  if (h.flags & F_SCALE) change_pref(r, s[S_SCALE]);
  entry_xform(r, h, pool);
  apertures(r, h, pool);
  if (h.fringe & 0xff) r = multipole_fringe(r, a, b, h.order, 1.0);
  r = exact_slices(r, a, b, h.order, h.nslice, s, rad);
  r.ct -= s[S_ELEN];  // absolute path length -> path lengthening
  if (h.fringe & 0xff00) r = multipole_fringe(r, a, b, h.order, -1.0);
  apertures(r, h, pool);
  exit_xform(r, h, pool);
  if (h.flags & F_SCALE) change_pref(r, s[S_INVSCALE]);
}

// CorrectorPass.c:25-90
__device__ __forceinline__ void corrector(R6& r, cDevElem& h, cdouble* pool) {
  cdouble* s = pool + h.sc;
  const double xk = s[S_C_XK], yk = s[S_C_YK];
  if (h.flags & F_SCALE) change_pref(r, s[S_SCALE]);
  if (!(h.flags & F_THICK)) {
    entry_xform(r, h, pool);
    apertures(r, h, pool);
    r.px += xk;
    r.py += yk;
  } else {
#if ATX_STRICT
    const double pn = 1.0 / (1.0 + r.dp);
#else
    const double pn = fast_rcp(1.0 + r.dp);
#endif
    const double nl = s[S_C_LEN] * pn;
    entry_xform(r, h, pool);
    apertures(r, h, pool);
    // s[S_C_Q0] = xk*xk/3 + yk*yk/3, the sum's first two terms
    r.ct += nl * pn * (s[S_C_Q0] + r.px * r.px + r.py * r.py + r.px * xk + r.py * yk) / 2.0;
    r.x += nl * (r.px + xk / 2.0);
    r.px += xk;
    r.y += nl * (r.py + yk / 2.0);
    r.py += yk;
    apertures(r, h, pool);
  }
  exit_xform(r, h, pool);
  if (h.flags & F_SCALE) change_pref(r, s[S_INVSCALE]);
}

#if ATX_AUX  // diffusion matrices: built with the reference operation order only
// --- Diffusion matrices -----------------------------------------------------
// Restatement of the reference's Ohmi diffusion-matrix propagation
// (diff_drift.c, diff_thinkick.c, diff_str_kick.c, diff_bnd_kick.c,
// diff_bend_fringe.c, diff_exactdrift.c, diff_str_exactkick.c and
// ATsandwichmmt/ATaddmm of atlalib.c:92-158), used with bdiff != NULL by
// the radiating passes. Matrices are column-major, [i][j] at [i + 6 j], and
// the products keep the reference summation order.

// B <- M B M^T (atlalib.c:129-147)
__device__ __forceinline__ void sandwich(const double* m, double* b) {
  double t[36];
#pragma unroll
  for (int i = 0; i < 6; i++) {
#pragma unroll
    for (int j = 0; j < 6; j++) {
      double c = 0;
#pragma unroll
      for (int k = 0; k < 6; k++) c += m[j + 6 * k] * b[k + 6 * i];
      t[j + 6 * i] = c;
    }
  }
#pragma unroll
  for (int i = 0; i < 6; i++) {
#pragma unroll
    for (int j = 0; j < 6; j++) {
      double c = 0;
#pragma unroll
      for (int k = 0; k < 6; k++) c += t[i + 6 * k] * m[j + 6 * k];
      b[i + 6 * j] = c;
    }
  }
}

__device__ __forceinline__ void eye6(double* m) {
#pragma unroll
  for (int i = 0; i < 36; i++) m[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++) m[7 * i] = 1.0;
}

// diff_drift.c:2-43
__device__ __forceinline__ void d_drift(R6& r, double L, double* b) {
  const double pn = 1.0 / (1.0 + r.dp);
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  const double nl = L * pn;
  double m[36];
  eye6(m);
  m[6] = nl;
  m[20] = nl;
  m[24] = -nl * xpr;
  m[26] = -nl * ypr;
  m[11] = nl * xpr;
  m[23] = nl * ypr;
  m[29] = -nl * (xpr * xpr + ypr * ypr);
  sandwich(m, b);
  r.x += nl * r.px;
  r.y += nl * r.py;
  r.ct += nl * pn * (r.px * r.px + r.py * r.py) / 2.0;
}

// diff_exactdrift.c:2-43
__device__ __forceinline__ void d_exact_drift(R6& r, double L, double* b) {
  const double pn = 1.0 / (1.0 + r.dp);
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  const double nl = L * pn / sqrt(1.0 - xpr * xpr - ypr * ypr);
  const double den = 1.0 - xpr * xpr - ypr * ypr;
  double m[36];
  eye6(m);
  m[6] = nl * (1.0 - ypr * ypr) / den;
  m[18] = nl * xpr * ypr / den;
  m[24] = -nl * xpr / den;
  m[8] = nl * xpr * ypr / den;
  m[20] = nl * (1.0 - xpr * xpr) / den;
  m[26] = -nl * ypr / den;
  m[11] = nl * xpr / den;
  m[23] = nl * ypr / den;
  m[29] = -nl * (xpr * xpr + ypr * ypr) / den;
  sandwich(m, b);
  r.x += nl * r.px;
  r.y += nl * r.py;
  r.ct += nl * (1.0 + r.dp);
}

// thinkickM of diff_thinkick.c:1-35 (irho = 0) and diff_bnd_kick.c:1-46
__device__ __forceinline__ void d_kick_m(const R6& r, cdouble* a, cdouble* bp, int n, double L,
                                         double irho, bool bend, double* b) {
  double imn = n * a[n];
  double ren = n * bp[n];
  for (int k = n - 1; k > 0; k--) {
    const double t = (ren * r.x - imn * r.y) + k * bp[k];
    imn = imn * r.x + ren * r.y + k * a[k];
    ren = t;
  }
  double m[36];
  eye6(m);
  m[1] = -L * ren;
  m[13] = L * imn;
  m[3] = L * imn;
  m[15] = L * ren;
  if (bend) {
    m[25] = L * irho;
    m[1] += -L * irho * irho;
    m[5] = L * irho;
  }
  sandwich(m, b);
}

// thinkickB (diff_thinkick.c:37-71, diff_bnd_kick.c:48-83)
__device__ __forceinline__ void d_kick_b(const R6& r, double difc, double b2p, double factor,
                                         double* b) {
  const double pn = 1.0 / (1.0 + r.dp);
  const double pn2 = pn * pn;
  const double b3p = b2p * sqrt(b2p);
  const double bb = difc * b3p * factor / pn2;
  double a[36];
#pragma unroll
  for (int i = 0; i < 36; i++) a[i] = 0.0;
  a[7] = bb * (r.px * r.px) * pn2;
  a[19] = bb * r.px * r.py * pn2;
  a[9] = a[19];
  a[21] = bb * (r.py * r.py) * pn2;
  a[10] = bb * r.px * pn;
  a[25] = a[10];
  a[22] = bb * r.py * pn;
  a[27] = a[22];
  a[28] = bb;
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = 0; j < 6; j++) b[i + 6 * j] += a[i + 6 * j];
}

// diff_str_kick.c:19-83 (bend = false), diff_bnd_kick.c:101-174 (bend = true)
__device__ __forceinline__ void d_rad_kick(R6& r, cdouble* a, cdouble* bp, int n, double L,
                                           double irho, double radc, double difc, bool bend,
                                           double* b) {
  double re = bp[n], im = a[n];
  for (int k = n - 1; k >= 0; k--) {
    const double t = re * r.x - im * r.y + bp[k];
    im = im * r.x + re * r.y + a[k];
    re = t;
  }
  double pn = 1.0 / (1.0 + r.dp);
  const double x = r.x;
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  double b2p, factor;
  if (bend) {
    const double by = re + irho, bx = im;
    const double hx = 1 + x * irho;
    const double vn2 = 1.0 / ((1.0 + x * irho) * (1.0 + x * irho) + xpr * xpr + ypr * ypr);
    b2p = ((by * hx) * (by * hx) + (bx * hx) * (bx * hx) + (bx * ypr - by * xpr) * (bx * ypr - by * xpr)) * vn2;
    factor = (1.0 + x * irho + (xpr * xpr + ypr * ypr) / 2.0) / (pn * pn) * L;
  } else {
    const double vn2 = 1.0 / (1.0 + xpr * xpr + ypr * ypr);
    const double cr = im * ypr - re * xpr;
    b2p = (re * re + im * im + cr * cr) * vn2;
    factor = (1.0 + (xpr * xpr + ypr * ypr) / 2.0) / (pn * pn) * L;
  }
  d_kick_m(r, a, bp, n, L, irho, bend, b);
  d_kick_b(r, difc, b2p, factor, b);
  const double d0 = r.dp;
  r.dp -= radc * b2p * factor;
  pn = 1.0 / (1.0 + r.dp);
  r.px = xpr / pn;
  r.py = ypr / pn;
  if (bend) {
    r.px -= L * (re - (d0 - r.x * irho) * irho);
    r.py += L * im;
    r.ct += L * irho * r.x;
  } else {
    r.px -= L * re;
    r.py += L * im;
  }
}

// diff_str_exactkick.c:14-71
__device__ __forceinline__ void d_exact_rad_kick(R6& r, cdouble* a, cdouble* bp, int n, double L,
                                                 double radc, double difc, double* b) {
  double re = bp[n], im = a[n];
  for (int k = n - 1; k >= 0; k--) {
    const double t = re * r.x - im * r.y + bp[k];
    im = im * r.x + re * r.y + a[k];
    re = t;
  }
  double pn = 1.0 / (1.0 + r.dp);
  const double xpr = r.px * pn;
  const double ypr = r.py * pn;
  const double t = im * xpr + re * ypr;
  const double b2p = im * im + re * re - t * t;
  const double factor = L / (pn * pn) / sqrt(1.0 - xpr * xpr - ypr * ypr);
  d_kick_m(r, a, bp, n, L, 0.0, false, b);
  d_kick_b(r, difc, b2p, factor, b);
  r.dp -= radc * b2p * factor;
  pn = 1.0 / (1.0 + r.dp);
  r.px = xpr / pn;
  r.py = ypr / pn;
  r.px -= L * re;
  r.py += L * im;
}

// diff_bend_fringe.c:2-53; mode and constants as for tracking
__device__ __forceinline__ void d_edge(R6& r, int mode, double irho, double fx, double c0,
                                       double edge_angle, double psi, double sgn, bool bterm,
                                       double* b) {
  const double pn = 1.0 / (1.0 + r.dp);
  double fy;
  if (mode == E_CONST)
    fy = c0;
  else if (mode == E_DIVP)
    fy = c0 / (1.0 + r.dp);
  else if (mode == E_MULPN)
    fy = c0 * pn;
  else
    fy = edge_fy_tan(mode, r.dp, r.px, irho, edge_angle, psi, sgn);
#pragma unroll
  for (int m = 0; m < 6; m++) {
    b[1 + 6 * m] += fx * b[6 * m];
    b[3 + 6 * m] -= fy * b[2 + 6 * m];
  }
  if (bterm)
#pragma unroll
    for (int m = 0; m < 6; m++)
      b[3 + 6 * m] -= b[4 + 6 * m] * r.y * (irho * irho + fy * fy) * psi * pn * pn / irho;
#pragma unroll
  for (int m = 0; m < 6; m++) {
    b[m + 6 * 1] += fx * b[m + 6 * 0];
    b[m + 6 * 3] -= fy * b[m + 6 * 2];
  }
  if (bterm)
#pragma unroll
    for (int m = 0; m < 6; m++)
      b[m + 6 * 3] -= b[m + 6 * 4] * r.y * (irho * irho + fy * fy) * psi * pn * pn / irho;
  r.px += r.x * fx;
  r.py -= r.y * fy;
}

// One radiating element from its entrance orbit: the path of
// StrMPoleSymplectic4RadPass.c:61-110, BndMPoleSymplectic4RadPass.c:76-123
// and ExactMultipoleRadPass.c:56-94 with the matrix updates.
__device__ __noinline__ void diffusion_elem(R6 r, cDevElem& h, cdouble* pool, double* b) {
  cdouble* s = pool + h.sc;
  cdouble* a = pool + h.poly;
  cdouble* bp = a + h.order + 1;
  const int n = h.order;
  const bool bend = h.kind == K_BNDRAD;
  const bool exact = h.kind == K_EXMPOLERAD;
  const double irho = s[S_IRHO];
  const double radc = s[S_RADC], difc = s[S_DIFC];
  const double l1 = s[S_L1], l2 = s[S_L2], k1 = s[S_K1], k2 = s[S_K2];
  if (h.flags & F_SCALE) change_pref(r, s[S_SCALE]);
  entry_xform(r, h, pool);
  apertures(r, h, pool);
  if (exact) {
    if (h.fringe & 0xff) r = multipole_fringe(r, a, bp, n, 1.0);
    for (int m = 0; m < h.nslice; m++) {
      d_exact_drift(r, l1, b);
      d_exact_rad_kick(r, a, bp, n, k1, radc, difc, b);
      d_exact_drift(r, l2, b);
      d_exact_rad_kick(r, a, bp, n, k2, radc, difc, b);
      d_exact_drift(r, l2, b);
      d_exact_rad_kick(r, a, bp, n, k1, radc, difc, b);
      d_exact_drift(r, l1, b);
    }
    return;  // the exit maps do not change B
  }
  if (bend)
    d_edge(r, h.edge & 0xff, irho, s[S_FX_IN], s[S_FY_IN], s[S_EDGE_IN], s[S_CORR_IN], 1.0,
           (h.flags & F_EDGEB_IN) != 0, b);
  if (h.fringe & 0xff) quad_fringe_in(r, h, pool, s[S_B2]);
  for (int m = 0; m < h.nslice; m++) {
    d_drift(r, l1, b);
    d_rad_kick(r, a, bp, n, k1, irho, radc, difc, bend, b);
    d_drift(r, l2, b);
    d_rad_kick(r, a, bp, n, k2, irho, radc, difc, bend, b);
    d_drift(r, l2, b);
    d_rad_kick(r, a, bp, n, k1, irho, radc, difc, bend, b);
    d_drift(r, l1, b);
  }
  if (h.fringe & 0xff00) quad_fringe_out(r, h, pool, s[S_B2]);
  if (bend)
    d_edge(r, (h.edge >> 8) & 0xff, irho, s[S_FX_OUT], s[S_FY_OUT], s[S_EDGE_OUT], s[S_CORR_OUT],
           -1.0, (h.flags & F_EDGEB_OUT) != 0, b);
}

__global__ void __launch_bounds__(64) diffusion_kernel(DiffArgs args) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= args.nelem) return;
  cDevElem& h = ((cDevElem*)args.elems)[e];
  double b[36];
#pragma unroll
  for (int i = 0; i < 36; i++) b[i] = 0.0;
  if (h.kind == K_STRRAD || h.kind == K_BNDRAD || h.kind == K_EXMPOLERAD) {
    const double* o = args.orbits + 6 * (int64_t)e;
    R6 r = {o[0], o[1], o[2], o[3], o[4], o[5]};
    if (!__builtin_isnan(r.x)) diffusion_elem(r, h, (cdouble*)args.pool, b);
  }
  double* out = args.bdiff + 36 * (int64_t)e;
#pragma unroll
  for (int i = 0; i < 36; i++) out[i] = b[i];
}

#endif  // ATX_AUX

// The kinds other than drifts and thick multipoles, for one particle. EXT:
// the kernel variant that also runs CavityPass and the exact passes;
// lattices without them use the variant without, whose element switch keeps
// the register demand of the common passes only.
template <bool EXT>
__device__ __forceinline__ void element_misc(R6& r, cDevElem& h, cdouble* pool, int nturn) {
  if (EXT) {
    switch (h.kind) {
      case K_CAVITY: r = cavity(r, pool + h.sc, (h.flags & F_THICK) != 0); return;
      case K_EXDRIFT:  // ExactDriftPass.c:15-44
        entry_xform(r, h, pool);
        apertures(r, h, pool);
        exact_drift(r, pool[h.sc + S_D_LEN]);
        r.ct -= pool[h.sc + S_D_LEN];
        apertures(r, h, pool);
        exit_xform(r, h, pool);
        return;
      case K_EXMPOLE: exact_mpole(r, h, pool, false); return;
      case K_EXMPOLERAD: exact_mpole(r, h, pool, true); return;
      default: break;
    }
  }
  switch (h.kind) {
    case K_IDENTITY:  // IdentityPass.c:20-43
      entry_xform(r, h, pool);
      apertures(r, h, pool);
      exit_xform(r, h, pool);
      break;
    case K_THIN: thin(r, h, pool); break;
    case K_RF: rf_cavity(r, h, pool, nturn); break;
    case K_CORR: corrector(r, h, pool); break;
    case K_APERTURE: rect_ap(r, pool + h.rap); break;  // AperturePass.c:9-26
    default: break;
  }
}

// The P particles of a lane through one element: drifts and thick
// multipoles step all P together, the other kinds run per particle.
// A drift's length, kept in the header words its kind does not use.
__device__ __forceinline__ double drift_length(cDevElem& h) {
  const uint64_t bits = ((uint64_t)(uint32_t)h.fringe << 32) | (uint32_t)h.edge;
  return __builtin_bit_cast(double, bits);
}

template <bool EXT, int P, int U>
__device__ __forceinline__ void element_lane(R6* r, cDevElem& h, cdouble* pool, int nturn) {
  if (h.kind == K_DRIFT) {  // DriftPass.c:16-48 with misalignment or apertures
    const double len = drift_length(h);
#pragma unroll
    for (int q = 0; q < P; q++) {
      entry_xform(r[q], h, pool);
      apertures(r[q], h, pool);
    }
#pragma unroll
    for (int q = 0; q < P; q++) drift6(r[q], len);
#pragma unroll
    for (int q = 0; q < P; q++) {
      apertures(r[q], h, pool);
      exit_xform(r[q], h, pool);
    }
    return;
  }
  switch (h.kind) {
    case K_STR: mpole_dispatch<false, false, P, U>(r, h, pool); break;
    case K_BND: mpole_dispatch<true, false, P, U>(r, h, pool); break;
    case K_STRRAD: mpole_dispatch<false, true, P, U>(r, h, pool); break;
    case K_BNDRAD: mpole_dispatch<true, true, P, U>(r, h, pool); break;
    default:
      element_misc<EXT>(r[0], h, pool, nturn);
      if constexpr (P > 1) element_misc<EXT>(r[1], h, pool, nturn);
      break;
  }
}

// at.c:115-139 predicate: any coordinate not finite, or |r_n| > 1 for n < 5.
__device__ __forceinline__ bool is_lost(const R6& r) {
  return !(__builtin_fabs(r.x) <= 1.0) || !(__builtin_fabs(r.px) <= 1.0) ||
         !(__builtin_fabs(r.y) <= 1.0) || !(__builtin_fabs(r.py) <= 1.0) ||
         !(__builtin_fabs(r.dp) <= 1.0) || !(__builtin_fabs(r.ct) < __builtin_inf());
}

__device__ __forceinline__ void store6(double* base, int64_t cs, const R6& r) {
  base[0] = r.x;
  base[cs] = r.px;
  base[2 * cs] = r.y;
  base[3 * cs] = r.py;
  base[4 * cs] = r.dp;
  base[5 * cs] = r.ct;
}

// Dynamic LDS of a workgroup of G particle slots, used to compact surviving
// particles into the leading lanes: 6*G doubles of state, G ints of particle
// index (bit 31: loss already recorded), 2 x 8 per-wave counts
// (double-buffered by round, so that a wave can publish its next count while
// a slower wave still reads the current ones).
__host__ __device__ constexpr size_t compact_lds_bytes(int slots) {
  return (size_t)slots * (6 * sizeof(double) + sizeof(int)) + 2 * 8 * sizeof(int);
}

__device__ __forceinline__ void store_state(cInstance& I, int p, const R6& r) {
  double* rp = I.r + (int64_t)p * I.pstride;
  const int64_t cs = I.cstride;
  rp[0] = r.x;
  rp[cs] = r.px;
  rp[2 * cs] = r.y;
  rp[3 * cs] = r.py;
  rp[4 * cs] = r.dp;
  rp[5 * cs] = r.ct;
}

__device__ __forceinline__ R6 lost_state() { return R6{__builtin_nan(""), 0.0, 0.0, 0.0, 0.0, 0.0}; }

// P particles per lane, lane-major: lane l of a workgroup holds its slots
// l*P .. l*P+P-1. live_begin (repacked turn blocks after the first; NULL
// otherwise): first workgroup of each instance among the workgroups that
// hold survivors, with live_begin[ninst] their total and live_begin[ninst +
// 1] the particle slots per workgroup (blockDim.x * P, or 64 * P when few
// survive), so that the live workgroups of all instances are the leading
// ones of the grid and the hardware spreads them evenly over the CUs; the
// other workgroups end at once.
template <bool EXT, int P, int U>
__global__ void __launch_bounds__(512) track_kernel(const Instance* __restrict__ insts_g, int ninst,
                                                    int nblocks, int xcd_order,
                                                    const int32_t* __restrict__ live_begin) {
  static_assert(P == 1 || P == 2, "one or two particles per lane");
  extern __shared__ double lds[];
  cInstance* insts = (cInstance*)insts_g;
  // Workgroup order (xcd_order: a launch over several instances of equal
  // turn counts): the hardware deals workgroup i to XCD i mod kXcds, and
  // each XCD has its own L2. The workgroups of the launch (nblocks, or the
  // live ones of a repacked turn block) are dealt in kXcds contiguous
  // chunks, one per XCD, so that the instances (lattices) of one chunk are
  // read through one L2 rather than every L2 holding every lattice. The
  // grid is kXcds * ceil(nblocks / kXcds) workgroups; the first kXcds *
  // ceil(nlog / kXcds) of them map one to one onto the logical workgroups
  // below that count, and those past nlog, like the rest, end at once. The
  // hardware order stands otherwise: one lattice is read alike by every
  // XCD, and items of unequal turn counts, dealt in chunks, would load the
  // XCDs unevenly.
  const int nlog = live_begin ? live_begin[ninst] : nblocks;
  const int hw = (int)blockIdx.x;
  int bid = hw;
  if (xcd_order) {
    const int chunk = (nlog + kXcds - 1) / kXcds;
    if (hw >= kXcds * chunk) return;
    bid = (hw % kXcds) * chunk + hw / kXcds;
  }
  if (bid >= nlog) return;
  // Instance of this workgroup: the last whose first block is <= bid.
  int lo = 0, hi = ninst - 1;
  int wbegin;
  int G = (int)blockDim.x * P;  // particle slots per workgroup
  if (live_begin) {
    G = live_begin[ninst + 1];
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (live_begin[mid] <= bid) lo = mid; else hi = mid - 1;
    }
    wbegin = live_begin[lo];
  } else {
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (insts[mid].block_begin <= bid) lo = mid; else hi = mid - 1;
    }
    wbegin = insts[lo].block_begin;
  }
  cInstance& I = insts[lo];
  const int tid = (int)threadIdx.x;
  const int npart = I.npart;
  // lanes of a repacked launch take the surviving particles idx[0..nlive)
  const int32_t* idx = I.idx;
  const int nslot = idx ? __builtin_amdgcn_readfirstlane(*I.nlive) : npart;
  const int first = (bid - wbegin) * G;  // first slot of this workgroup
  // particles of this workgroup, and its waves that hold some (uniform);
  // the other waves end now
  int held = max(0, min(G, nslot - first));
  int nw = (held + 64 * P - 1) / (64 * P);
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (wave >= nw) return;

  cDevElem* elems = (cDevElem*)I.elems;
  cdouble* pool = (cdouble*)I.pool;
  cint* slots = (cint*)I.slots;
  const int64_t cs = I.cstride;
  const int losses = I.losses;
  int p[P];
  bool has[P], lost[P];
  R6 r[P];
#pragma unroll
  for (int q = 0; q < P; q++) {
    const int sl = tid * P + q;
    has[q] = sl < held;  // this slot holds a particle
    p[q] = has[q] ? (idx ? idx[first + sl] : first + sl) : 0;
    r[q] = lost_state();
    if (has[q]) {
      const double* rp = I.r + (int64_t)p[q] * I.pstride;
      r[q].x = rp[0];
      r[q].px = rp[cs];
      r[q].y = rp[2 * cs];
      r[q].py = rp[3 * cs];
      r[q].dp = rp[4 * cs];
      r[q].ct = rp[5 * cs];
    }
    lost[q] = !has[q];  // slots without a particle never record a loss
    if (has[q]) {
      if (losses == 2) {
        lost[q] = I.lost_flag[p[q]] != 0;  // recorded by an earlier call
      } else if (losses) {
        I.lost_flag[p[q]] = 0;
        I.lost_turn[p[q]] = 0;
        I.lost_elem[p[q]] = 0;
        store6(I.lost_coord + 6 * (int64_t)p[q], 1, R6{0.0, 0.0, 0.0, 0.0, 0.0, 0.0});
      }
    }
  }

  const int nelem = I.nelem;
  const bool check = !I.elempass;
  double* rout = I.rout;
  const int64_t np6 = 6 * (int64_t)npart;
  const int nturns = I.nturns;
  const bool refs = (I.n_first | I.n_std) != 0;  // uniform: skip the slot reads without refpts
  // Lost particles are dropped at turn boundaries and the survivors packed
  // into the leading lanes: with one particle per lane when packing frees a
  // wave (the waves left without particles end and free their SIMD slots),
  // with several whenever one was lost (a lane steps its particles together
  // only while none of them is lost). Not with refpts: lost particles still
  // fill their refpt copies, as in at.c.
  const bool compact = check && !refs && nturns > 1;
  double* s_r = lds;
  int* s_p = (int*)(lds + 6 * G);
  int* s_cnt = s_p + G;
  for (int t = 0; t < nturns; t++) {
    if (compact && t > 0) {
      // A particle whose x is NaN is lost for good (elements skip it and the
      // loss check leaves it as it is): store its final state now.
      int mine = 0, below = 0;
      int within[P];
#pragma unroll
      for (int q = 0; q < P; q++) {
        if (has[q] && __builtin_isnan(r[q].x)) {
          store_state(I, p[q], r[q]);
          has[q] = false;
        }
        const uint64_t m = __ballot(has[q]);
        within[q] = mine;
        mine += has[q] ? 1 : 0;
        below += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
      }
      int cnt = mine;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
      int* cnt_buf = s_cnt + 8 * (t & 1);
      if ((tid & 63) == 0) cnt_buf[wave] = cnt;
      __syncthreads();
      int off = 0, total = 0;
      for (int k = 0; k < nw; k++) {
        const int c = cnt_buf[k];
        off += (k < wave) ? c : 0;
        total += c;
      }
      total = __builtin_amdgcn_readfirstlane(total);
      const int need = (total + 64 * P - 1) / (64 * P);
      if (need < nw || (P > 1 && total < held)) {  // uniform
#pragma unroll
        for (int q = 0; q < P; q++) {
          if (has[q]) {
            const int j = off + below + within[q];
            s_r[j] = r[q].x;
            s_r[G + j] = r[q].px;
            s_r[2 * G + j] = r[q].y;
            s_r[3 * G + j] = r[q].py;
            s_r[4 * G + j] = r[q].dp;
            s_r[5 * G + j] = r[q].ct;
            s_p[j] = p[q] | (lost[q] ? (int)0x80000000u : 0);
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < P; q++) {
          const int j = tid * P + q;
          has[q] = j < total;
          if (has[q]) {
            r[q].x = s_r[j];
            r[q].px = s_r[G + j];
            r[q].y = s_r[2 * G + j];
            r[q].py = s_r[3 * G + j];
            r[q].dp = s_r[4 * G + j];
            r[q].ct = s_r[5 * G + j];
            const int v = s_p[j];
            p[q] = v & 0x7fffffff;
            lost[q] = v < 0;
          } else {
            r[q] = lost_state();
            lost[q] = true;
          }
        }
        held = total;
        nw = need;
        if (wave >= nw) return;  // this wave holds no particle any more
      }
    }
    const int nturn = I.turn0 + t;
    const int e0 = (t == 0) ? I.start_elem : 0;
    cint* slot = slots + ((t == 0) ? 0 : nelem + 1);
    const int64_t base = (t == 0) ? 0 : (int64_t)I.n_first + (int64_t)(t - 1) * I.n_std;
    const int e1 = I.end_elem;
    for (int e = e0; e < e1; e++) {
      if (refs) {
        const int sl = slot[e];
#pragma unroll
        for (int q = 0; q < P; q++)
          if (has[q] && sl >= 0) store6(rout + (base + sl) * np6 + 6 * (int64_t)p[q], 1, r[q]);
      }
      cDevElem& h = elems[e];
      bool unchanged = false;  // no coordinate changed: the loss check would repeat the last
      if (h.kind == K_DRIFT && h.flags == 0) {
        // A drift without misalignment or apertures (DriftPass.c:16-48), half
        // the elements of a ring: tested first, from the header's first two
        // words, and stepped without the element dispatch. Every element
        // skips a particle whose x is NaN (at.c:142-161).
        const double len = drift_length(h);
#pragma unroll
        for (int q = 0; q < P; q++) {
#if ATX_STRICT
          if (!__builtin_isnan(r[q].x)) drift6(r[q], len);
#else
          // the fast build skips a NaN-x lane by a zero step instead of
          // selecting its old coordinates: x stays NaN and a finite y, ct
          // stay as they are (a lost particle's are zero). The operations
          // and their rounding are drift6's in this build.
          const double pn = fast_rcp(1.0 + r[q].dp);
          const double nl = __builtin_isnan(r[q].x) ? 0.0 : len * pn;
          r[q].x = __builtin_fma(nl, r[q].px, r[q].x);
          r[q].y = __builtin_fma(nl, r[q].py, r[q].y);
          const double p2 = __builtin_fma(r[q].px, r[q].px, r[q].py * r[q].py);
          const double dct = (pn * nl) * p2;
          const double hct = dct * 0.5;
          r[q].ct = r[q].ct + hct;
#endif
        }
      } else if (h.kind == K_IDENTITY && h.flags == 0) {
        // A marker or monitor without misalignment or apertures
        // (IdentityPass.c:20-43) changes no coordinate, so its loss check
        // would repeat the one after the element before it: only the first
        // element of the call checks.
        unchanged = t > 0 || e > e0;
      } else if (P == 1) {
        if (!__builtin_isnan(r[0].x)) element_lane<EXT, 1, U>(r, h, pool, nturn);
      } else {
        // A lane steps its particles together. Every element skips a
        // particle whose x is NaN (at.c:142-161, *Pass.c): in a wave that
        // holds such a particle (a loss since the last compaction, or a NaN
        // input), its state is kept in the lane's own LDS slots and put back
        // after the element. Slots without a particle step too; they are
        // never stored.
        bool stale[P];
        bool any = false;
#pragma unroll
        for (int q = 0; q < P; q++) {
          stale[q] = has[q] && __builtin_isnan(r[q].x);
          any |= stale[q];
        }
        const bool keep = __ballot(any) != 0;  // uniform
        if (keep) {
#pragma unroll
          for (int q = 0; q < P; q++) {
            const int j = tid * P + q;
            if (stale[q]) {
              s_r[j] = r[q].x;
              s_r[G + j] = r[q].px;
              s_r[2 * G + j] = r[q].y;
              s_r[3 * G + j] = r[q].py;
              s_r[4 * G + j] = r[q].dp;
              s_r[5 * G + j] = r[q].ct;
            }
          }
        }
        element_lane<EXT, P, U>(r, h, pool, nturn);
        if (keep) {
#pragma unroll
          for (int q = 0; q < P; q++) {
            const int j = tid * P + q;
            if (stale[q]) {
              r[q].x = s_r[j];
              r[q].px = s_r[G + j];
              r[q].y = s_r[2 * G + j];
              r[q].py = s_r[3 * G + j];
              r[q].dp = s_r[4 * G + j];
              r[q].ct = s_r[5 * G + j];
            }
          }
        }
      }
      if (check && !unchanged) {
#pragma unroll
        for (int q = 0; q < P; q++) {
          if (losses) {
            if (!lost[q] && is_lost(r[q])) {
              lost[q] = true;
              I.lost_flag[p[q]] = 1;
              I.lost_turn[p[q]] = (uint32_t)nturn;
              I.lost_elem[p[q]] = (uint32_t)e;
              store6(I.lost_coord + 6 * (int64_t)p[q], 1, r[q]);
              r[q] = lost_state();
            }
          } else if (has[q] && __builtin_isfinite(r[q].x) && is_lost(r[q])) {
            r[q] = lost_state();  // at.c:142-161
          }
        }
      }
    }
    if (refs) {
      const int sl = slot[nelem];
#pragma unroll
      for (int q = 0; q < P; q++)
        if (has[q] && sl >= 0) store6(rout + (base + sl) * np6 + 6 * (int64_t)p[q], 1, r[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < P; q++)
    if (has[q]) store_state(I, p[q], r[q]);
}

}  // namespace ATX_NS
}  // namespace atx

// The kernel of this object's launch shape: nblocks workgroups of `block`
// lanes, block * ATX_PPT particle slots each.
extern "C" hipError_t ATX_LAUNCH(const atx::Instance* d_insts, int ninst, int nblocks, int block,
                                 int ext, int xcd_order, const int32_t* live_begin,
                                 hipStream_t stream) {
  const size_t lds = atx::ATX_NS::compact_lds_bytes(block * ATX_PPT);
  // whole rounds of XCDs: the kernel's workgroup order needs the grid to
  // cover kXcds * ceil(n / kXcds) for every n <= nblocks
  const int grid = (nblocks + atx::kXcds - 1) / atx::kXcds * atx::kXcds;
  if (ext)
    hipLaunchKernelGGL((atx::ATX_NS::track_kernel<true, ATX_PPT, ATX_UNROLL>), dim3(grid),
                       dim3(block), lds, stream, d_insts, ninst, nblocks, xcd_order, live_begin);
  else
    hipLaunchKernelGGL((atx::ATX_NS::track_kernel<false, ATX_PPT, ATX_UNROLL>), dim3(grid),
                       dim3(block), lds, stream, d_insts, ninst, nblocks, xcd_order, live_begin);
  return hipGetLastError();
}

#if ATX_AUX
namespace atx {
// Probes of atx_element_matrices (matrix.py:195-200): particle 12 e + c is
// the entrance orbit of element e plus column c of [D, -D], D = diag(step/2):
// +step/2 on coordinate c (c < 6), -step/2 on coordinate c - 6 (c >= 6), and
// +0.0 (c < 6) or -0.0 (c >= 6) elsewhere, as numpy adds them; (6, 12 nelem)
// F-order.
__global__ void __launch_bounds__(256) element_probe_kernel(const double* __restrict__ orbits,
                                                            double step, double* __restrict__ r,
                                                            int nelem) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 12 * (int64_t)nelem) return;
  const int64_t e = i / 12;
  const int c = (int)(i % 12);
  const double d = (c < 6) ? 0.5 * step : -(0.5 * step);
  const double z = (c < 6) ? 0.0 : -0.0;
#pragma unroll
  for (int k = 0; k < 6; k++) r[6 * i + k] = orbits[6 * e + k] + ((k == c % 6) ? d : z);
}

// m[i][j] = (probe j+ - probe j-)_i / step (matrix.py:201-202), column-major
// 6x6 per element ([i][j] at 36 e + i + 6 j).
__global__ void __launch_bounds__(256) element_diff_kernel(const double* __restrict__ r,
                                                           double step, double* __restrict__ m66,
                                                           int nelem) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 36 * (int64_t)nelem) return;
  const int64_t e = t / 36;
  const int i = (int)(t % 6), j = (int)((t % 36) / 6);
  const double* p = r + 72 * e;
  m66[t] = (p[6 * j + i] - p[6 * (6 + j) + i]) / step;
}

// Repack between the turn blocks of a batched launch: one workgroup per
// instance lists, in order, the particles still alive (x not NaN: a lost
// particle is set to NaN, at.c:142-161) in idx[0..*nlive). The first pass
// scans 0..npart, later passes the previous list, in place: a chunk is read
// completely before it is written, and writes never pass the read front.
__global__ void __launch_bounds__(1024) repack_kernel(const Instance* __restrict__ insts,
                                                      int first_pass) {
  __shared__ int s_wave[16];
  const Instance& I = insts[blockIdx.x];
  int32_t* idx = const_cast<int32_t*>(I.idx);
  int32_t* nlive = const_cast<int32_t*>(I.nlive);
  if (!idx) return;
  const int n = first_pass ? I.npart : *nlive;
  const int tid = (int)threadIdx.x;
  const int wave = tid >> 6;
  const int nwaves = (int)blockDim.x >> 6;
  int out = 0;
  for (int base = 0; base < n; base += (int)blockDim.x) {
    const int s = base + tid;
    int q = 0;
    bool alive = false;
    if (s < n) {
      q = first_pass ? s : idx[s];
      alive = !__builtin_isnan(I.r[(int64_t)q * I.pstride]);
    }
    const uint64_t m = __ballot(alive);
    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                               __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
    if ((tid & 63) == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    int off = 0, total = 0;
    for (int k = 0; k < nwaves; k++) {
      const int c = s_wave[k];
      off += (k < wave) ? c : 0;
      total += c;
    }
    if (alive) idx[out + off + rank] = q;
    out += total;
    __syncthreads();
  }
  if (tid == 0) *nlive = out;
}
}  // namespace atx

namespace atx {
// Live particles of instance i with turns in the block (0 for the others).
__device__ __forceinline__ int live_count(const Instance& I) {
  return I.nturns > 0 ? (I.idx ? *I.nlive : I.npart) : 0;
}

// Map of the live workgroups for the next turn block (one workgroup):
// live_begin[i] = sum over j < i of ceil(nlive_j / g), live_begin[ninst] =
// the total and live_begin[ninst + 1] = g, the particles per workgroup: the
// block size, or 64 when all survivors fit in fine_limit lanes (then each
// workgroup runs one wave and the dispatcher spreads single waves over the
// SIMDs, which a tail of few survivors needs; the grid holds enough
// workgroups for either). Chunks of 1,024 instances: a wave-level scan, then
// the wave totals through LDS, carrying the running sum from chunk to chunk.
__global__ void __launch_bounds__(1024) live_map_kernel(const Instance* __restrict__ insts, int ninst,
                                                        int block, int fine, int fine_limit,
                                                        int32_t* __restrict__ live_begin) {
  __shared__ int s_wave[16];
  __shared__ int s_total;
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = (int)blockDim.x >> 6;
  // total survivors
  int part = 0;
  for (int i = tid; i < ninst; i += (int)blockDim.x) part += live_count(insts[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if (lane == 0) s_wave[wave] = part;
  __syncthreads();
  if (tid == 0) {
    int t = 0;
    for (int k = 0; k < nwaves; k++) t += s_wave[k];
    s_total = t;
  }
  __syncthreads();
  const int g = s_total <= fine_limit ? fine : block;
  int carry = 0;
  for (int base = 0; base < ninst; base += (int)blockDim.x) {
    const int i = base + tid;
    int c = 0;
    if (i < ninst) c = (live_count(insts[i]) + g - 1) / g;
    int v = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int u = __shfl_up(v, off, 64);
      if (lane >= off) v += u;
    }
    if (lane == 63) s_wave[wave] = v;
    __syncthreads();
    int woff = 0, total = 0;
    for (int k = 0; k < nwaves; k++) {
      const int w = s_wave[k];
      woff += (k < wave) ? w : 0;
      total += w;
    }
    if (i < ninst) live_begin[i] = carry + woff + v - c;
    carry += total;
    __syncthreads();
  }
  if (tid == 0) {
    live_begin[ninst] = carry;
    live_begin[ninst + 1] = g;
  }
}
}  // namespace atx

extern "C" hipError_t atx_launch_repack(const atx::Instance* d_insts, int ninst, int first_pass,
                                        hipStream_t stream) {
  if (ninst <= 0) return hipSuccess;
  hipLaunchKernelGGL(atx::repack_kernel, dim3(ninst), dim3(1024), 0, stream, d_insts, first_pass);
  return hipGetLastError();
}

extern "C" hipError_t atx_launch_live_map(const atx::Instance* d_insts, int ninst, int block,
                                          int fine, int fine_limit, int32_t* live_begin,
                                          hipStream_t stream) {
  if (ninst <= 0) return hipSuccess;
  hipLaunchKernelGGL(atx::live_map_kernel, dim3(1), dim3(1024), 0, stream, d_insts, ninst, block,
                     fine, fine_limit, live_begin);
  return hipGetLastError();
}

extern "C" hipError_t atx_launch_element_probes(const double* orbits, double step, double* r,
                                                int nelem, hipStream_t stream) {
  const int64_t n = 12 * (int64_t)nelem;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(atx::element_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     stream, orbits, step, r, nelem);
  return hipGetLastError();
}

extern "C" hipError_t atx_launch_element_diffs(const double* r, double step, double* m66,
                                               int nelem, hipStream_t stream) {
  const int64_t n = 36 * (int64_t)nelem;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(atx::element_diff_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     stream, r, step, m66, nelem);
  return hipGetLastError();
}

extern "C" hipError_t atx_launch_diffusion(const atx::DiffArgs* args, hipStream_t stream) {
  const int block = 64;
  const int nblocks = (args->nelem + block - 1) / block;
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(atx::ATX_NS::diffusion_kernel, dim3(nblocks), dim3(block), 0, stream, *args);
  return hipGetLastError();
}
#endif
