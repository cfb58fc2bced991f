// atx_adjoint.hip -- reverse-mode (adjoint) kernel of the tracking map.
//
// Gradient of a loss of the final particle coordinates with respect to the
// state at the start of a block of turns and to the multipole coefficients
// PolynomB[k] of selected elements. One lane per particle walks the turns of
// the block backwards and, inside each turn, the elements and their
// Forest-Ruth sub-steps in reverse order. The input state of every sub-step
// is reconstructed by inverting that sub-step exactly in closed form (every
// map of the supported pass methods is a shear: drift, thin kick, edge
// kick, RF energy kick, corrector, misalignment), starting each turn from
// the exact turn-start state recorded by the forward tracking kernel, so no
// per-element storage is needed and reconstruction error never spans more
// than one turn. The adjoint uses the transposed per-step Jacobians of the
// reference tangent maps (diff_drift.c:2-28, diff_thinkick.c:1-35,
// diff_bnd_kick.c:1-46, edge terms of diff_bend_fringe.c:32-48).
#include <hip/hip_runtime.h>

#include "atx_device.h"
#include "atx_libm.h"

namespace atx {
namespace adj {

typedef __attribute__((address_space(4))) const double cdouble;
typedef __attribute__((address_space(4))) const DevElem cDevElem;
typedef __attribute__((address_space(4))) const AdjElem cAdjElem;
typedef __attribute__((address_space(4))) const AdjItem cAdjItem;

struct V6 {
  double x, px, y, py, dp, ct;
};

// polynomial value and complex derivative: sum (B_k + i A_k) z^k, z = x + i y
template <int N>
__device__ __forceinline__ void field_d(double x, double y, cdouble* a, cdouble* b, int n,
                                        double& re, double& im, double& ren, double& imn) {
  const int order = (N >= 0) ? N : n;
  re = b[order];
  im = a[order];
  ren = 0.0;
  imn = 0.0;
#pragma unroll
  for (int k = order - 1; k >= 0; k--) {
    const double tr = ren * x - imn * y + re;
    imn = ren * y + imn * x + im;
    ren = tr;
    const double t = re * x - im * y + b[k];
    im = re * y + im * x + a[k];
    re = t;
  }
}

// Re, Im of z^j for the coefficient gradient
__device__ __forceinline__ void zpow(double x, double y, int j, double& zr, double& zi) {
  zr = 1.0;
  zi = 0.0;
  for (int i = 0; i < j; i++) {
    const double t = zr * x - zi * y;
    zi = zr * y + zi * x;
    zr = t;
  }
}

// Reverse of an in-magnet drift x' = x + nl px, ct' = ct + nl (px^2+py^2) u/2
// with nl = L u, u = 1/(1+dp).
__device__ __forceinline__ void drift_rev(V6& r, V6& l, double len) {
  const double u = 1.0 / (1.0 + r.dp);
  const double nl = len * u;
  const double p2 = r.px * r.px + r.py * r.py;
  r.x -= nl * r.px;
  r.y -= nl * r.py;
  r.ct -= nl * p2 * 0.5 * u;
  const double nlu = nl * u;
  l.dp -= nlu * (r.px * l.x + r.py * l.y + p2 * u * l.ct);
  l.px += nl * l.x + nlu * r.px * l.ct;
  l.py += nl * l.y + nlu * r.py * l.ct;
}

// Reverse of a thin multipole kick of integrated length k (straight or
// curved frame). g accumulates dL/dB_jg.
template <bool BEND, int N>
__device__ __forceinline__ void kick_rev(V6& r, V6& l, cdouble* a, cdouble* b, int n, double k,
                                         double h, int jg, double& g) {
  double re, im, ren, imn;
  field_d<N>(r.x, r.y, a, b, n, re, im, ren, imn);
  if (jg >= 0) {
    double zr, zi;
    zpow(r.x, r.y, jg, zr, zi);
    g += k * (zi * l.py - zr * l.px);
  }
  const double lx = -k * ren * l.px + k * imn * l.py;
  const double ly = k * imn * l.px + k * ren * l.py;
  if (BEND) {
    l.x += lx - k * h * h * l.px + k * h * l.ct;
    l.dp += k * h * l.px;
    r.px += k * (re - (r.dp - r.x * h) * h);
    r.ct -= k * h * r.x;
  } else {
    l.x += lx;
    r.px += k * re;
  }
  l.y += ly;
  r.py -= k * im;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ void grad_flush(double g, int slot, double* grad) {
  const double s = wave_sum(g);
  if ((threadIdx.x & 63) == 0) atomicAdd(grad + slot, s);
}

// misalignment: state r <- Minv r, adjoint l <- M^T l
__device__ __noinline__ void mat_rev(V6& r, V6& l, cdouble* m, cdouble* minv) {
  const double v[6] = {r.x, r.px, r.y, r.py, r.dp, r.ct};
  const double w[6] = {l.x, l.px, l.y, l.py, l.dp, l.ct};
  double o[6], q[6];
  for (int i = 0; i < 6; i++) {
    double s = 0.0, t = 0.0;
    for (int j = 0; j < 6; j++) {
      s += minv[i + 6 * j] * v[j];
      t += m[j + 6 * i] * w[j];
    }
    o[i] = s;
    q[i] = t;
  }
  r = V6{o[0], o[1], o[2], o[3], o[4], o[5]};
  l = V6{q[0], q[1], q[2], q[3], q[4], q[5]};
}

__device__ __forceinline__ void sub_t(V6& r, cdouble* t) {
  r.x -= t[0];
  r.px -= t[1];
  r.y -= t[2];
  r.py -= t[3];
  r.dp -= t[4];
  r.ct -= t[5];
}

__device__ __forceinline__ void exit_rev(V6& r, V6& l, cDevElem& h, cAdjElem& ae, cdouble* pool) {
  if (h.flags & F_T2) sub_t(r, pool + h.t2);
  if (h.flags & F_R2) mat_rev(r, l, pool + h.r2, pool + ae.r2inv);
}

__device__ __forceinline__ void entry_rev(V6& r, V6& l, cDevElem& h, cAdjElem& ae, cdouble* pool) {
  if (h.flags & F_R1) mat_rev(r, l, pool + h.r1, pool + ae.r1inv);
  if (h.flags & F_T1) sub_t(r, pool + h.t1);
}

// edge kick px' = px + x fx, py' = py - y fy (E_CONST) or fy = c0/(1+dp) (E_DIVP)
__device__ __forceinline__ void edge_rev(V6& r, V6& l, int mode, double fx, double c0) {
  const double u = 1.0 / (1.0 + r.dp);
  const double fy = (mode == E_DIVP) ? c0 * u : c0;
  if (mode == E_DIVP) l.dp += r.y * c0 * u * u * l.py;
  l.x += fx * l.px;
  l.y -= fy * l.py;
  r.px -= r.x * fx;
  r.py += r.y * fy;
}

template <bool BEND, int N>
__device__ __forceinline__ void mpole_rev(V6& r, V6& l, cDevElem& h, cAdjElem& ae, cdouble* pool,
                                          double* grad) {
  cdouble* s = pool + h.sc;
  cdouble* a = pool + h.poly;
  cdouble* b = a + h.order + 1;
  const int n = h.order;
  const int jg = (ae.slot >= 0 && grad) ? ae.coeff : -1;
  double g = 0.0;
  const double irho = s[S_IRHO];
  exit_rev(r, l, h, ae, pool);
  if (BEND) edge_rev(r, l, (h.edge >> 8) & 0xff, s[S_FX_OUT], s[S_FY_OUT]);
  const double l1 = s[S_L1], l2 = s[S_L2], k1 = s[S_K1], k2 = s[S_K2];
  for (int m = 0; m < h.nslice; m++) {
    drift_rev(r, l, l1);
    kick_rev<BEND, N>(r, l, a, b, n, k1, irho, jg, g);
    drift_rev(r, l, l2);
    kick_rev<BEND, N>(r, l, a, b, n, k2, irho, jg, g);
    drift_rev(r, l, l2);
    kick_rev<BEND, N>(r, l, a, b, n, k1, irho, jg, g);
    drift_rev(r, l, l1);
  }
  if (BEND) edge_rev(r, l, h.edge & 0xff, s[S_FX_IN], s[S_FY_IN]);
  entry_rev(r, l, h, ae, pool);
  if (jg >= 0) grad_flush(g, ae.slot, grad);
}

template <bool BEND>
__device__ __forceinline__ void mpole_rev_dispatch(V6& r, V6& l, cDevElem& h, cAdjElem& ae,
                                                   cdouble* pool, double* grad) {
  switch (h.order) {
    case 1: mpole_rev<BEND, 1>(r, l, h, ae, pool, grad); break;
    case 2: mpole_rev<BEND, 2>(r, l, h, ae, pool, grad); break;
    case 3: mpole_rev<BEND, 3>(r, l, h, ae, pool, grad); break;
    default: mpole_rev<BEND, -1>(r, l, h, ae, pool, grad); break;
  }
}

__device__ __forceinline__ void thin_rev(V6& r, V6& l, cDevElem& h, cAdjElem& ae, cdouble* pool,
                                         double* grad) {
  cdouble* s = pool + h.sc;
  cdouble* a = pool + h.poly;
  cdouble* b = a + h.order + 1;
  const int jg = (ae.slot >= 0 && grad) ? ae.coeff : -1;
  double g = 0.0;
  exit_rev(r, l, h, ae, pool);
  const double bax = s[S_BAX], bay = s[S_BAY];
  // ct' = ct - bax x + bay y; px' = px + bax dp; py' = py - bay dp
  r.ct += bax * r.x - bay * r.y;
  r.px -= bax * r.dp;
  r.py += bay * r.dp;
  l.x -= bax * l.ct;
  l.y += bay * l.ct;
  l.dp += bax * l.px - bay * l.py;
  kick_rev<false, -1>(r, l, a, b, h.order, 1.0, 0.0, jg, g);
  entry_rev(r, l, h, ae, pool);
  if (jg >= 0) grad_flush(g, ae.slot, grad);
}

// full drift with its own 1/(1+dp): same Jacobian as drift_rev
__device__ __forceinline__ void rf_rev(V6& r, V6& l, cDevElem& h, cdouble* pool, int nturn) {
  cdouble* s = pool + h.sc;
  const double nt = (double)nturn;
  if (h.flags & F_THICK) drift_rev(r, l, s[S_RF_HALF]);
  if (h.flags & F_RFKICK) {
    const double tw = s[S_RF_TWOPIF];
    double ph, sn;
    {
      // the forward kick's phase and sine, bit for bit, so the inverse is exact
#pragma clang fp contract(off)
      ph = tw * ((r.ct - s[S_RF_LAG]) / 2.99792458e8 - s[S_RF_HFT0] * nt) - s[S_RF_PHILAG];
      sn = libm::host_sin(ph);
    }
    const double cs = cos(ph);
    r.dp += s[S_RF_NV] * sn;  // undo dp += -nv sin(ph)
    l.ct += -s[S_RF_NV] * cs * tw / 2.99792458e8 * l.dp;
  }
  if (h.flags & F_THICK) drift_rev(r, l, s[S_RF_HALF]);
}

__device__ __forceinline__ void corrector_rev(V6& r, V6& l, cDevElem& h, cAdjElem& ae,
                                              cdouble* pool) {
  cdouble* s = pool + h.sc;
  const double xk = s[S_C_XK], yk = s[S_C_YK];
  exit_rev(r, l, h, ae, pool);
  r.px -= xk;
  r.py -= yk;
  if (h.flags & F_THICK) {
    // forward (from the input state): ct += nl u Q/2, x += nl (px + xk/2),
    // y += nl (py + yk/2), with nl = L u, u = 1/(1+dp)
    const double u = 1.0 / (1.0 + r.dp);
    const double nl = s[S_C_LEN] * u;
    r.x -= nl * (r.px + 0.5 * xk);
    r.y -= nl * (r.py + 0.5 * yk);
    const double q = xk * xk / 3.0 + yk * yk / 3.0 + r.px * r.px + r.py * r.py + r.px * xk + r.py * yk;
    r.ct -= nl * u * q * 0.5;
    const double nlu = nl * u;
    l.dp -= nlu * ((r.px + 0.5 * xk) * l.x + (r.py + 0.5 * yk) * l.y + q * u * l.ct);
    l.px += nl * l.x + nlu * (r.px + 0.5 * xk) * l.ct;
    l.py += nl * l.y + nlu * (r.py + 0.5 * yk) * l.ct;
  }
  entry_rev(r, l, h, ae, pool);
}

__device__ __forceinline__ void element_rev(V6& r, V6& l, cDevElem& h, cAdjElem& ae, cdouble* pool,
                                            double* grad, int nturn) {
  switch (h.kind) {
    case K_IDENTITY:
      exit_rev(r, l, h, ae, pool);
      entry_rev(r, l, h, ae, pool);
      break;
    case K_DRIFT:
      exit_rev(r, l, h, ae, pool);
      drift_rev(r, l, pool[h.sc + S_D_LEN]);
      entry_rev(r, l, h, ae, pool);
      break;
    case K_STR: mpole_rev_dispatch<false>(r, l, h, ae, pool, grad); break;
    case K_BND: mpole_rev_dispatch<true>(r, l, h, ae, pool, grad); break;
    case K_THIN: thin_rev(r, l, h, ae, pool, grad); break;
    case K_RF: rf_rev(r, l, h, pool, nturn); break;
    case K_CORR: corrector_rev(r, l, h, ae, pool); break;
    default: break;  // AperturePass: identity for surviving particles
  }
}

// grid: x over the particles of the largest item, y over the items; lanes
// beyond an item's particle count idle with zero adjoint (a workgroup with
// none of its particles returns at once: the kernel has no barrier)
__global__ void __launch_bounds__(256) adjoint_kernel(AdjArgs args) {
  cAdjItem& it = ((cAdjItem*)args.items)[blockIdx.y];
  const int npart = it.npart;
  if ((int)(blockIdx.x * blockDim.x) >= npart) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < npart;
  const int64_t np6 = 6 * (int64_t)npart;
  const double* states = it.states;
  double* grad = it.grad;
  cDevElem* elems = (cDevElem*)args.elems;
  cAdjElem* aes = (cAdjElem*)args.adj;
  cdouble* pool = (cdouble*)args.pool;
  // inactive lanes (beyond N, or particles lost by the end of the block)
  // run with zero adjoint so that every lane reaches the wave reductions
  const double* send = states + (int64_t)it.nturns * np6 + 6 * (int64_t)(active ? p : 0);
  bool live = active && __builtin_isfinite(send[0]);
  V6 l = {0, 0, 0, 0, 0, 0};
  if (live) {
    const double* lp = it.lam + 6 * (int64_t)p;
    l = V6{lp[0], lp[1], lp[2], lp[3], lp[4], lp[5]};
  }
  for (int t = it.nturns - 1; t >= 0; t--) {
    const double* sp = states + (int64_t)(t + 1) * np6 + 6 * (int64_t)(active ? p : 0);
    V6 r = {0, 0, 0, 0, 0, 0};
    if (live) r = V6{sp[0], sp[1], sp[2], sp[3], sp[4], sp[5]};
    const int nturn = it.turn0 + t;
    for (int e = args.nelem - 1; e >= 0; e--) element_rev(r, l, elems[e], aes[e], pool, grad, nturn);
  }
  if (active) {
    double* lp = it.lam + 6 * (int64_t)p;
    lp[0] = l.x;
    lp[1] = l.px;
    lp[2] = l.y;
    lp[3] = l.py;
    lp[4] = l.dp;
    lp[5] = l.ct;
  }
}

}  // namespace adj
}  // namespace atx

// max_npart: the largest particle count of the items (host-side copy of the
// device item table); nitems <= 65535 (grid y limit)
extern "C" hipError_t atx_launch_adjoint(const atx::AdjArgs* args, int32_t max_npart,
                                         hipStream_t stream) {
  const int block = 256;
  const int nblocks = (max_npart + block - 1) / block;
  if (nblocks == 0 || args->nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(atx::adj::adjoint_kernel, dim3(nblocks, args->nitems), dim3(block), 0, stream,
                     *args);
  return hipGetLastError();
}
