// atx_host.cpp -- host runtime of the tracking engine: element attribute
// store, lattice compiler (validation + host precompute into the device
// format of atx_device.h), device memory, launch, and the C ABI of
// include/atx.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/atx.h"
#include "atx_device.h"

// Tracking kernels, one per build and launch shape (particles per lane x
// slices per loop iteration), compiled from atx_kernel.hip as separate objects.
#define ATX_TRACK_DECL(name)                                                                  \
  extern "C" hipError_t name(const atx::Instance*, int, int, int, int, int, const int32_t*,       \
                             hipStream_t);
ATX_TRACK_DECL(atx_launch_track_strict_p1_u1)
ATX_TRACK_DECL(atx_launch_track_strict_p1_u2)
ATX_TRACK_DECL(atx_launch_track_strict_p1_u4)
ATX_TRACK_DECL(atx_launch_track_strict_p2_u1)
ATX_TRACK_DECL(atx_launch_track_strict_p2_u2)
ATX_TRACK_DECL(atx_launch_track_fast_p1_u1)
ATX_TRACK_DECL(atx_launch_track_fast_p1_u2)
ATX_TRACK_DECL(atx_launch_track_fast_p1_u4)
ATX_TRACK_DECL(atx_launch_track_fast_p2_u1)
ATX_TRACK_DECL(atx_launch_track_fast_p2_u2)
#undef ATX_TRACK_DECL
extern "C" hipError_t atx_launch_diffusion(const atx::DiffArgs*, hipStream_t);
extern "C" hipError_t atx_launch_element_probes(const double*, double, double*, int, hipStream_t);
extern "C" hipError_t atx_launch_element_diffs(const double*, double, double*, int, hipStream_t);
extern "C" hipError_t atx_launch_repack(const atx::Instance*, int, int, hipStream_t);
extern "C" hipError_t atx_launch_live_map(const atx::Instance*, int, int, int, int, int32_t*,
                                          hipStream_t);
extern "C" hipError_t atx_launch_adjoint(const atx::AdjArgs*, int32_t, hipStream_t);

namespace atx {

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Constants of atintegrators/atconstants.h:15-31 (values, not code).
const double kCbrt2 = 1.25992104989487316477;
const double kDrift1 = 1.0 / (2.0 * (2.0 - kCbrt2));
const double kDrift2 = 0.5 - (1.0 / (2.0 * (2.0 - kCbrt2)));
const double kKick1 = 1.0 / (2.0 - kCbrt2);
const double kKick2 = 1.0 - 2.0 / (2.0 - kCbrt2);
const double kC0 = 2.99792458e8;
const double kTwoPi = 2.0 * 3.14159265358979323846;
const double kRe = 2.8179403205e-15;
const double kE0GeV = 510.99895069e-6;
const double kRadConst = 2.0 / 3.0 * kRe;
// atconstants.h:24-31
const double kDifConst = 55.0 / 24.0 / 1.7320508075688772 * 197.3269804e-18 / kE0GeV * kRe;

struct Attr {
  std::vector<double> v;
  bool is_int = false;
};

struct PassInfo {
  const char* name;
  int kind;
  bool mpole;
};

// Pass methods of the engine (names of atgpu/PassMethodFactory.cpp:33-46
// restricted to the in-scope ones).
const PassInfo kPasses[] = {
    {"IdentityPass", K_IDENTITY, false},
    {"DriftPass", K_DRIFT, false},
    {"BndMPoleSymplectic4Pass", K_BND, true},
    {"BndMPoleSymplectic4RadPass", K_BNDRAD, true},
    {"StrMPoleSymplectic4Pass", K_STR, true},
    {"StrMPoleSymplectic4RadPass", K_STRRAD, true},
    {"RFCavityPass", K_RF, false},
    {"ThinMPolePass", K_THIN, true},
    {"CorrectorPass", K_CORR, false},
    {"AperturePass", K_APERTURE, false},
    {"CavityPass", K_CAVITY, false},
    {"ExactDriftPass", K_EXDRIFT, false},
    {"ExactMultipolePass", K_EXMPOLE, true},
    {"ExactMultipoleRadPass", K_EXMPOLERAD, true},
};

struct HostElem {
  std::string pass;
  std::string fam;
  int kind = 0;
  std::map<std::string, Attr> attrs;

  const Attr* get(const char* n) const {
    auto it = attrs.find(n);
    return it == attrs.end() ? nullptr : &it->second;
  }
};

struct Compiled {
  std::vector<DevElem> elems;
  std::vector<double> pool;
  double T0 = 0.0;
  std::vector<int32_t> r1inv, r2inv;  // per element, pool offsets (-1: none)
  std::vector<std::string> adjerr;    // per element, why the adjoint cannot run it ("" = it can)
};

// Gauss-Jordan inverse of a 6x6 column-major matrix with partial pivoting.
bool invert6(const double* m, double* inv) {
  double a[6][12];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      a[i][j] = m[i + 6 * j];
      a[i][j + 6] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 6; c++) {
    int p = c;
    for (int i = c + 1; i < 6; i++)
      if (std::fabs(a[i][c]) > std::fabs(a[p][c])) p = i;
    if (a[p][c] == 0.0) return false;
    if (p != c)
      for (int j = 0; j < 12; j++) std::swap(a[p][j], a[c][j]);
    const double d = a[c][c];
    for (int j = 0; j < 12; j++) a[c][j] /= d;
    for (int i = 0; i < 6; i++) {
      if (i == c) continue;
      const double f = a[i][c];
      for (int j = 0; j < 12; j++) a[i][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) inv[i + 6 * j] = a[i][j + 6];
  return true;
}

// row and column 4 (delta) of the matrix are those of the identity
bool delta_preserving(const double* m) {
  for (int k = 0; k < 6; k++) {
    const double e = (k == 4) ? 1.0 : 0.0;
    if (m[4 + 6 * k] != e || m[k + 6 * 4] != e) return false;
  }
  return true;
}

// Drift (c) and kick (d) coefficients per slice of the integrator types of
// atgpu/SymplecticIntegrator.cpp:32-86 (published coefficient sets: Euler;
// McLachlan & Atela 1991 2nd and 4th order; Ruth 1983 3rd order; Forest &
// Ruth 1990 4th order; Yoshida 1990 6th order, solution A).
struct Scheme {
  int n;
  double c[8], d[8];
};

bool integrator_scheme(int type, Scheme* s) {
  const double r2 = 1.0 / std::sqrt(2.0);
  switch (type) {
    case 1: *s = {1, {1.0}, {1.0}}; return true;
    case 2: *s = {2, {1.0 - r2, r2}, {1.0 - r2, r2}}; return true;
    case 3: *s = {3, {1.0, -2.0 / 3.0, 2.0 / 3.0}, {-1.0 / 24.0, 3.0 / 4.0, 7.0 / 24.0}}; return true;
    case 4: *s = {4, {kDrift1, kDrift2, kDrift2, kDrift1}, {kKick1, kKick2, kKick1, 0.0}}; return true;
    case 5:
      *s = {4, {0.1288461583653843, 0.4415830236164665, -0.0857820194129737, 0.5153528374311229},
               {0.3340036032863214, 0.7563200005156683, -0.2248198030794208, 0.1344961992774311}};
      return true;
    case 6:  // the 16-digit values atgpu tabulates from Yoshida's w1..w3
      *s = {8,
            {0.3922568052387800, 0.5100434119184585, -0.4710533854097566, 0.0687531682525181,
             0.0687531682525181, -0.4710533854097566, 0.5100434119184585, 0.3922568052387800},
            {0.7845136104775600, 0.2355732133593570, -1.1776799841788701, 1.3151863206839063,
             -1.1776799841788701, 0.2355732133593570, 0.7845136104775600, 0.0}};
      return true;
    default:
      return false;
  }
}

}  // namespace
}  // namespace atx

struct atx_lattice {
  std::vector<atx::HostElem> elems;
  // device copy of the compiled lattice, valid for (energy, rest, elempass)
  bool dev_valid = false;
  bool has_ext = false;  // CavityPass or exact passes present: extended kernel variant
  double dev_energy = -1.0, dev_rest = -1.0;
  bool dev_elempass = false;
  int dev_device = -1;
  atx::DevElem* d_elems = nullptr;
  double* d_pool = nullptr;
  size_t d_elems_n = 0, d_pool_n = 0;
  int64_t turn_counter = 0;
  int integrator = 4;
  atx_launch_shape shape{0, 0, 0, 0};  // atx_track / atx_elempass launches (0: default)
  // host copies of per-element compile results used by the adjoint
  std::vector<int32_t> h_r1inv, h_r2inv, h_order, h_kind;
  std::vector<std::string> h_adjerr;
  std::mutex mu;
};

namespace atx {
namespace {

// --- attribute access with the reference's required/optional semantics ----

struct Reader {
  const HostElem& e;
  int idx;
  int err = 0;

  void missing(const char* n) {
    if (!err)
      err = fail(ATX_E_ATTRIBUTE, std::string("The required attribute ") + n +
                                      " is missing (element #" + std::to_string(idx) + " " +
                                      e.fam + ", " + e.pass + ")");
  }
  void bad(const std::string& m) {
    if (!err) err = fail(ATX_E_VALUE, m + " (element #" + std::to_string(idx) + " " + e.fam + ")");
  }
  double dbl(const char* n) {
    const Attr* a = e.get(n);
    if (!a || a->v.empty()) {
      missing(n);
      return 0.0;
    }
    return a->v[0];
  }
  double opt_dbl(const char* n, double d) {
    const Attr* a = e.get(n);
    return (a && !a->v.empty()) ? a->v[0] : d;
  }
  long lng(const char* n) { return (long)dbl(n); }
  long opt_lng(const char* n, long d) {
    const Attr* a = e.get(n);
    return (a && !a->v.empty()) ? (long)a->v[0] : d;
  }
  const std::vector<double>* arr(const char* n) {
    const Attr* a = e.get(n);
    if (!a) {
      missing(n);
      return nullptr;
    }
    return &a->v;
  }
  const std::vector<double>* opt_arr(const char* n, size_t need) {
    const Attr* a = e.get(n);
    if (!a) return nullptr;
    if (need && a->v.size() < need) {
      bad(std::string("attribute ") + n + " has " + std::to_string(a->v.size()) +
          " values, expected " + std::to_string(need));
      return nullptr;
    }
    return &a->v;
  }
};

int32_t push(std::vector<double>& pool, const double* v, size_t n) {
  int32_t off = (int32_t)pool.size();
  pool.insert(pool.end(), v, v + n);
  return off;
}

// Common optional misalignment and aperture attributes.
void transforms(Reader& rd, DevElem& h, std::vector<double>& pool) {
  if (auto t = rd.opt_arr("T1", 6)) { h.flags |= F_T1; h.t1 = push(pool, t->data(), 6); }
  if (auto t = rd.opt_arr("R1", 36)) { h.flags |= F_R1; h.r1 = push(pool, t->data(), 36); }
  if (auto t = rd.opt_arr("R2", 36)) { h.flags |= F_R2; h.r2 = push(pool, t->data(), 36); }
  if (auto t = rd.opt_arr("T2", 6)) { h.flags |= F_T2; h.t2 = push(pool, t->data(), 6); }
  if (auto t = rd.opt_arr("RApertures", 4)) { h.flags |= F_RAP; h.rap = push(pool, t->data(), 4); }
  if (auto t = rd.opt_arr("EApertures", 2)) { h.flags |= F_EAP; h.eap = push(pool, t->data(), 2); }
}

// elegant linear fringe blocks (quadfringe.c:52-93) for one part.
void elegant_block(double k1, const double* fi, int part, bool swap, double out[8]) {
  const double k1sq = k1 * k1;
  double j1x, j2x, j3x, j1y;
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
  const double j2y = -j2x, j3y = j3x;
  const double ex = std::exp(j1x), ey = std::exp(j1y);
  double r00 = ex, r01 = j2x / ex, r10 = ex * j3x, r11 = (1 + j2x * j3x) / ex;
  double r22 = ey, r23 = j2y / ey, r32 = ey * j3y, r33 = (1 + j2y * j3y) / ey;
  if (swap) {
    std::swap(r00, r11);
    std::swap(r22, r33);
  }
  out[0] = r00; out[1] = r01; out[2] = r10; out[3] = r11;
  out[4] = r22; out[5] = r23; out[6] = r32; out[7] = r33;
}

// Edge-focusing mode and constants for one side of a bend
// (atphyslib.c:15-79 non-radiating, diff_bend_fringe.c:2-53 radiating).
int32_t edge_mode(bool rad, int method, double irho, double angle, double fint, double gap,
                  double* fx, double* c0, double* corr) {
  double cr;
  if ((fint == 0.0) || (gap == 0.0) || (method == 0)) {
    cr = 0.0;
  } else {
    const double se = std::sin(angle), ce = std::cos(angle);
    cr = rad ? irho * gap * fint * (1.0 + se * se) / ce : irho * gap * fint * (1 + se * se) / ce;
  }
  *fx = irho * std::tan(angle);
  *c0 = *fx;
  *corr = cr;
  if (method == 3) return rad ? E_RTAN3 : E_TAN3;
  if (method == 2) {
    if (cr == 0.0) return rad ? E_MULPN : E_DIVP;
    return rad ? E_RTAN2 : E_TAN2;
  }
  if (cr == 0.0) return E_CONST;
  return rad ? E_RTAN1 : E_TAN1;
}

int compile_elem(const HostElem& e, int idx, double energy, double rest, double T0, int integrator,
                 Compiled& out) {
  Reader rd{e, idx};
  DevElem h;
  std::memset(&h, 0, sizeof(h));
  h.kind = e.kind;
  h.ops = -1;
  std::vector<double>& pool = out.pool;
  double sc[S_NSCALAR];
  for (double& v : sc) v = 0.0;

  switch (e.kind) {
    case K_IDENTITY:
      transforms(rd, h, pool);
      break;
    case K_DRIFT:
    case K_EXDRIFT:  // DriftPass.c:50-80, ExactDriftPass.c:46-75
      sc[S_D_LEN] = rd.dbl("Length");
      if (e.kind == K_DRIFT) {
        // the tracking kernel reads a drift's length from its header (edge,
        // fringe: unused by drifts), saving a dependent load per drift
        uint64_t bits;
        std::memcpy(&bits, &sc[S_D_LEN], sizeof(bits));
        h.edge = (int32_t)(uint32_t)bits;
        h.fringe = (int32_t)(uint32_t)(bits >> 32);
      }
      transforms(rd, h, pool);
      break;
    case K_CAVITY: {  // CavityPass.c:62-88
      const double len = rd.dbl("Length");
      const double volt = rd.dbl("Voltage");
      const double freq = rd.dbl("Frequency");
      const double elem_energy = rd.opt_dbl("Energy", energy);
      const double lag = rd.opt_dbl("TimeLag", 0.0);
      const double philag = rd.opt_dbl("PhaseLag", 0.0);
      if (rd.err) break;
      sc[S_RF_HALF] = len / 2;
      sc[S_RF_NV] = volt / (energy == 0.0 ? elem_energy : energy);
      sc[S_RF_TWOPIF] = kTwoPi * freq;
      sc[S_RF_LAG] = lag;
      sc[S_RF_PHILAG] = philag;
      if (len != 0.0) h.flags |= F_THICK;
    } break;
    case K_APERTURE: {
      auto lim = rd.arr("Limits");
      if (lim && lim->size() < 4) rd.bad("Limits must have 4 values");
      if (!rd.err) h.rap = push(pool, lim->data(), 4);
    } break;
    case K_CORR: {
      const double len = rd.dbl("Length");
      auto ka = rd.arr("KickAngle");
      if (ka && ka->size() < 2) rd.bad("KickAngle must have 2 values");
      const double s = rd.opt_dbl("FieldScaling", 1.0);
      transforms(rd, h, pool);
      if (rd.err) break;
      sc[S_C_XK] = (*ka)[0];
      sc[S_C_YK] = (*ka)[1];
      sc[S_C_LEN] = len;
      // the constant leading terms of the thick corrector's path-length sum
      // (CorrectorPass.c:55-89), added first there, in the same order
      sc[S_C_Q0] = sc[S_C_XK] * sc[S_C_XK] / 3.0 + sc[S_C_YK] * sc[S_C_YK] / 3.0;
      sc[S_SCALE] = s;
      sc[S_INVSCALE] = 1.0 / s;
      if (s != 1.0) h.flags |= F_SCALE;
      if (len != 0.0) h.flags |= F_THICK;
    } break;
    case K_RF: {
      const double len = rd.dbl("Length");
      const double volt = rd.dbl("Voltage");
      const double freq = rd.dbl("Frequency");
      const double lag = rd.opt_dbl("TimeLag", 0.0);
      const double philag = rd.opt_dbl("PhaseLag", 0.0);
      if (rd.err) break;
      if (!(energy > 0.0)) {
        rd.err = fail(ATX_E_VALUE, "Energy must be positive. Check lattice or set the track() keyword argument.");
        break;
      }
      const double harm = std::round(freq * T0);  // RFCavityPass.c:59
      const double nv = volt / energy;            // RFCavityPass.c:65
      sc[S_RF_HALF] = len / 2;
      sc[S_RF_NV] = nv;
      sc[S_RF_TWOPIF] = kTwoPi * freq;
      sc[S_RF_LAG] = lag;
      sc[S_RF_PHILAG] = philag;
      sc[S_RF_HFT0] = harm / freq - T0;
      if (len != 0.0) h.flags |= F_THICK;
      if (nv != 0.0) h.flags |= F_RFKICK;
    } break;
    default: {  // multipoles
      const bool thin = (e.kind == K_THIN);
      const bool bend = (e.kind == K_BND || e.kind == K_BNDRAD);
      const bool exact = (e.kind == K_EXMPOLE || e.kind == K_EXMPOLERAD);
      const bool rad = (e.kind == K_STRRAD || e.kind == K_BNDRAD || e.kind == K_EXMPOLERAD);
      const double len = thin ? 0.0 : rd.dbl("Length");
      auto pa = rd.arr("PolynomA");
      auto pb = rd.arr("PolynomB");
      const long order = rd.lng("MaxOrder");
      const long nslice = thin ? 0 : rd.lng("NumIntSteps");
      double bang = 0.0, ein = 0.0, eout = 0.0;
      if (bend) {
        bang = rd.dbl("BendingAngle");
        ein = rd.dbl("EntranceAngle");
        eout = rd.dbl("ExitAngle");
      }
      if (rd.err) break;
      if (order < 0 || order > 30) {
        rd.bad("MaxOrder out of range");
        break;
      }
      if (exact && nslice <= 0) {  // ExactMultipolePass.c:126-128
        rd.err = fail(ATX_E_VALUE, "NumIntSteps must be positive");
        break;
      }
      const size_t need = (size_t)order + 1 > (thin ? 1u : 2u) ? (size_t)order + 1 : (thin ? 1u : 2u);
      if (pa->size() < (size_t)order + 1 || pb->size() < need) {
        rd.bad("PolynomA/PolynomB shorter than MaxOrder+1");
        break;
      }
      const double s = rd.opt_dbl("FieldScaling", 1.0);
      auto ka = rd.opt_arr("KickAngle", 2);
      transforms(rd, h, pool);
      if (rd.err) break;
      h.order = (int32_t)order;
      h.nslice = (int32_t)nslice;
      std::vector<double> a(pa->begin(), pa->begin() + order + 1);
      std::vector<double> b(pb->begin(), pb->begin() + order + 1);
      if (ka) {
        if (thin) {  // ThinMPolePass.c:38-41
          b[0] -= (*ka)[0];
          a[0] += (*ka)[1];
        } else {  // StrMPoleSymplectic4Pass.c:49-52
          b[0] -= std::sin((*ka)[0]) / len;
          a[0] += std::sin((*ka)[1]) / len;
        }
      }
      h.poly = push(pool, a.data(), a.size());
      push(pool, b.data(), b.size());
      sc[S_SCALE] = s;
      sc[S_INVSCALE] = 1.0 / s;
      if (s != 1.0) h.flags |= F_SCALE;
      if (thin) {
        if (const Attr* ba = e.get("BendingAngle")) {
          if (ba->v.size() >= 1) sc[S_BAX] = ba->v[0];
          if (ba->v.size() >= 2) sc[S_BAY] = ba->v[1];
        }
        break;
      }
      const double sl = len / (double)nslice;
      sc[S_L1] = sl * kDrift1;
      sc[S_L2] = sl * kDrift2;
      sc[S_HL1] = 0.5 * sc[S_L1];
      sc[S_HL2] = 0.5 * sc[S_L2];
      sc[S_K1] = sl * kKick1;
      sc[S_K2] = sl * kKick2;
      sc[S_ELEN] = len;
      if (exact) {
        // PTC multipole fringes of every order whenever the flag is set
        // (ExactMultipolePass.c:62-63,83-84); the Forest-Ruth scheme of the
        // CPU pass whatever the integrator type
        const long fqi = rd.opt_lng("FringeQuadEntrance", 0);
        const long fqo = rd.opt_lng("FringeQuadExit", 0);
        if (rd.err) break;
        h.fringe = (fqi ? 1 : 0) | ((fqo ? 1 : 0) << 8);
        if (rad) {
          if (!(energy > 0.0)) {
            rd.err = fail(ATX_E_VALUE, "Energy must be positive. Check lattice or set the track() keyword argument.");
            break;
          }
          const double gamma = (rest == 0.0) ? 1.0E-9 * energy / kE0GeV : energy / rest;
          sc[S_RADC] = kRadConst * std::pow(gamma, 3);
          sc[S_DIFC] = kDifConst * std::pow(gamma, 5);
        }
        break;
      }
      if (integrator != 4) {
        // op table: drifts and kicks of SL*c_i / SL*d_i, zero steps omitted
        // as atgpu's generated slice loop omits them (SymplecticIntegrator.cpp:140-171)
        Scheme sch;
        integrator_scheme(integrator, &sch);
        std::vector<double> ops(1, 0.0);
        for (int i = 0; i < sch.n; i++) {
          if (sch.c[i] != 0.0) { ops.push_back(0.0); ops.push_back(sl * sch.c[i]); }
          if (sch.d[i] != 0.0) { ops.push_back(1.0); ops.push_back(sl * sch.d[i]); }
        }
        ops[0] = (double)((ops.size() - 1) / 2);
        h.ops = push(pool, ops.data(), ops.size());
      }
      const double b2 = (*pb)[1];
      sc[S_B2] = b2;
      const long fqi = rd.opt_lng("FringeQuadEntrance", 0);
      const long fqo = rd.opt_lng("FringeQuadExit", 0);
      auto fim = rd.opt_arr("fringeIntM0", 5);
      auto fip = rd.opt_arr("fringeIntP0", 5);
      if (rd.err) break;
      const int qin = (fqi && b2 != 0) ? ((fqi == 2 && fim && fip) ? Q_ELEGANT : Q_LEE) : Q_NONE;
      const int qout = (fqo && b2 != 0) ? ((fqo == 2 && fim && fip) ? Q_ELEGANT : Q_LEE) : Q_NONE;
      h.fringe = qin | (qout << 8);
      if (qin == Q_ELEGANT || qout == Q_ELEGANT) {
        double m[32];
        elegant_block(b2, fip->data(), 2, true, m);        // entrance, first matrix
        elegant_block(b2, fim->data(), 1, true, m + 8);    // entrance, second matrix
        elegant_block(b2, fim->data(), 1, false, m + 16);  // exit, first matrix
        elegant_block(b2, fip->data(), 2, false, m + 24);  // exit, second matrix
        h.fq = push(pool, m, 32);
      }
      if (bend) {
        const double irho = bang / len;
        sc[S_IRHO] = irho;
        const int mi = (int)rd.opt_lng("FringeBendEntrance", 1);
        const int mo = (int)rd.opt_lng("FringeBendExit", 1);
        const double gap = rd.opt_dbl("FullGap", 0.0);
        const double f1 = rd.opt_dbl("FringeInt1", 0.0);
        const double f2 = rd.opt_dbl("FringeInt2", 0.0);
        const int32_t ein_mode = edge_mode(rad, mi, irho, ein, f1, gap, &sc[S_FX_IN], &sc[S_FY_IN], &sc[S_CORR_IN]);
        sc[S_EDGE_IN] = ein;
        const int32_t eout_mode = edge_mode(rad, mo, irho, eout, f2, gap, &sc[S_FX_OUT], &sc[S_FY_OUT], &sc[S_CORR_OUT]);
        sc[S_EDGE_OUT] = eout;
        h.edge = ein_mode | (eout_mode << 8);
        if (f1 > 0 && gap > 0) h.flags |= F_EDGEB_IN;
        if (f2 > 0 && gap > 0) h.flags |= F_EDGEB_OUT;
      }
      if (rad) {
        if (!(energy > 0.0)) {
          rd.err = fail(ATX_E_VALUE, "Energy must be positive. Check lattice or set the track() keyword argument.");
          break;
        }
        const double gamma = (rest == 0.0) ? 1.0E-9 * energy / kE0GeV : energy / rest;
        sc[S_RADC] = kRadConst * std::pow(gamma, 3);
        sc[S_DIFC] = kDifConst * std::pow(gamma, 5);
        sc[S_RK1] = sc[S_RADC] * sc[S_K1];
        sc[S_RK2] = sc[S_RADC] * sc[S_K2];
      }
    } break;
  }
  if (rd.err) return rd.err;
  h.sc = push(pool, sc, S_NSCALAR);
  // adjoint support: inverse misalignment matrices, and whether the reverse
  // kernel can invert this element in closed form
  int32_t r1i = -1, r2i = -1;
  std::string aerr;
  if (h.flags & F_R1) {
    double inv[36];
    if (!invert6(&pool[h.r1], inv)) aerr = "singular R1";
    r1i = push(pool, inv, 36);
    if (!delta_preserving(&pool[h.r1])) aerr = "R1 couples to delta";
  }
  if (h.flags & F_R2) {
    double inv[36];
    if (!invert6(&pool[h.r2], inv)) aerr = "singular R2";
    r2i = push(pool, inv, 36);
    if (!delta_preserving(&pool[h.r2])) aerr = "R2 couples to delta";
  }
  if ((h.flags & F_T1) && pool[h.t1 + 4] != 0.0) aerr = "T1 shifts delta";
  if ((h.flags & F_T2) && pool[h.t2 + 4] != 0.0) aerr = "T2 shifts delta";
  if (h.kind == K_STRRAD || h.kind == K_BNDRAD) aerr = "radiating pass";
  if (h.kind == K_CAVITY || h.kind == K_EXDRIFT || h.kind == K_EXMPOLE || h.kind == K_EXMPOLERAD)
    aerr = "pass method without an adjoint";
  if (h.flags & F_SCALE) aerr = "FieldScaling != 1";
  if (h.kind != K_DRIFT && h.fringe) aerr = "quadrupole fringe";  // a drift's holds its length
  if (h.ops >= 0) aerr = "integrator other than 4";
  for (int side = 0; side < 2; side++) {
    const int m = (h.edge >> (8 * side)) & 0xff;
    if ((h.kind == K_BND) && m != E_CONST && m != E_DIVP) aerr = "edge fringe with gap correction";
  }
  out.r1inv.push_back(r1i);
  out.r2inv.push_back(r2i);
  out.adjerr.push_back(aerr);
  out.elems.push_back(h);
  return ATX_OK;
}

int compile_lattice(const atx_lattice* lat, double energy, double rest, bool elempass,
                    Compiled& out) {
  double ring_length = 0.0;
  for (const auto& e : lat->elems) {
    const Attr* a = e.get("Length");
    if (a && !a->v.empty()) ring_length += a->v[0];
  }
  out.T0 = elempass ? 0.0 : ring_length / kC0;  // at.c:540-551, at.c:728-729
  out.elems.clear();
  out.pool.clear();
  for (size_t i = 0; i < lat->elems.size(); i++) {
    int rc = compile_elem(lat->elems[i], (int)i, energy, rest, out.T0, lat->integrator, out);
    if (rc) return rc;
  }
  if (out.pool.empty()) out.pool.push_back(0.0);
  return ATX_OK;
}

// Refpt copy ranks following at.c:618-651 (the output pointer advances
// once per copy, so a copy's slot is its rank in the copy sequence).
int turn_slots(const uint32_t* refpts, int nref, int start, int end, int nelem, int32_t* slot) {
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

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      return fail(ATX_E_RUNTIME, std::string(#x) + ": " + hipGetErrorString(e_));          \
  } while (0)

int ensure_device_lattice(atx_lattice* lat, double energy, double rest, bool elempass) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (lat->dev_valid && lat->dev_energy == energy && lat->dev_rest == rest &&
      lat->dev_elempass == elempass && lat->dev_device == dev)
    return ATX_OK;
  Compiled c;
  int rc = compile_lattice(lat, energy, rest, elempass, c);
  if (rc) return rc;
  if (lat->d_elems) (void)hipFree(lat->d_elems);
  if (lat->d_pool) (void)hipFree(lat->d_pool);
  lat->d_elems = nullptr;
  lat->d_pool = nullptr;
  lat->dev_valid = false;
  const size_t ne = c.elems.empty() ? 1 : c.elems.size();
  HIPCHK(hipMalloc(&lat->d_elems, ne * sizeof(DevElem)));
  HIPCHK(hipMalloc(&lat->d_pool, c.pool.size() * sizeof(double)));
  if (!c.elems.empty())
    HIPCHK(hipMemcpy(lat->d_elems, c.elems.data(), c.elems.size() * sizeof(DevElem), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(lat->d_pool, c.pool.data(), c.pool.size() * sizeof(double), hipMemcpyHostToDevice));
  lat->d_elems_n = c.elems.size();
  lat->d_pool_n = c.pool.size();
  lat->h_r1inv = c.r1inv;
  lat->h_r2inv = c.r2inv;
  lat->h_adjerr = c.adjerr;
  lat->h_order.clear();
  lat->h_kind.clear();
  lat->has_ext = false;
  for (const auto& h : c.elems) {
    lat->h_order.push_back(h.order);
    lat->h_kind.push_back(h.kind);
    lat->has_ext |= h.kind >= K_CAVITY;
  }
  lat->dev_valid = true;
  lat->dev_energy = energy;
  lat->dev_rest = rest;
  lat->dev_elempass = elempass;
  lat->dev_device = dev;
  return ATX_OK;
}

constexpr int kBlock = 256;
constexpr int kPpt = 1;
// two slices per loop iteration in the fast build (120 VGPRs, four waves per
// SIMD); one in the strict build, whose two-slice kernel needs 136 VGPRs and
// would run three waves per SIMD
constexpr int kUnrollFast = 2;
constexpr int kUnrollStrict = 1;

typedef hipError_t (*TrackLaunch)(const Instance*, int, int, int, int, int, const int32_t*,
                                  hipStream_t);

// XCD-aware workgroup order (the kernel deals the launch's workgroups to the
// XCDs in contiguous chunks, one L2 per group of instances): for launches
// over several instances that all run the same number of turns. Items of
// unequal turn counts keep the hardware order, which spreads every item
// over all XCDs.
bool xcd_order(const Instance* insts, size_t n) {
  if (n < 2) return false;
  for (size_t i = 1; i < n; i++)
    if (insts[i].nturns != insts[0].nturns) return false;
  return true;
}

// The tracking kernel of a build and a (valid) launch shape.
TrackLaunch track_launcher(bool strict, const atx_launch_shape& sh) {
  static const TrackLaunch table[2][2][3] = {
      {{atx_launch_track_fast_p1_u1, atx_launch_track_fast_p1_u2, atx_launch_track_fast_p1_u4},
       {atx_launch_track_fast_p2_u1, atx_launch_track_fast_p2_u2, nullptr}},
      {{atx_launch_track_strict_p1_u1, atx_launch_track_strict_p1_u2, atx_launch_track_strict_p1_u4},
       {atx_launch_track_strict_p2_u1, atx_launch_track_strict_p2_u2, nullptr}}};
  const int u = sh.unroll == 4 ? 2 : sh.unroll - 1;
  return table[strict ? 1 : 0][sh.ppt - 1][u];
}

// Launch shape with the defaults filled in, or an error for an invalid one.
int resolve_shape(const atx_launch_shape* in, atx_launch_shape& out, bool strict) {
  out = atx_launch_shape{kBlock, kPpt, strict ? kUnrollStrict : kUnrollFast, 0};
  if (in) {
    if (in->block) out.block = in->block;
    if (in->ppt) out.ppt = in->ppt;
    if (in->unroll) out.unroll = in->unroll;
  }
  if (out.block % 64 || out.block < 64 || out.block > 512)
    return fail(ATX_E_VALUE, "block size must be a multiple of 64 in [64, 512]");
  if (out.ppt != 1 && out.ppt != 2) return fail(ATX_E_VALUE, "particles per lane must be 1 or 2");
  if (out.unroll != 1 && out.unroll != 2 && out.unroll != 4)
    return fail(ATX_E_VALUE, "slice unroll must be 1, 2 or 4");
  // four slices of two particles need more than the 256 registers of a lane
  if (out.ppt == 2 && out.unroll == 4)
    return fail(ATX_E_VALUE, "two particles per lane take a slice unroll of 1 or 2");
  return ATX_OK;
}

// Pinned staging for the launch tables (instance tables, refpt slot tables,
// adjoint tables). hipMemcpyAsync from pageable memory returns only once the
// work queued before it on the stream has run, which would hold the host
// (and the streams it feeds next) behind every running kernel; from pinned
// memory it returns at once. A staging buffer is reused when the event
// recorded after its last copy has completed; buffers live until the
// process ends (at most one per upload in flight at a time).
struct Staging {
  char* p = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;  // recorded on streams of `device` only
  int device = 0;
};
std::mutex g_stage_mu;
std::vector<std::unique_ptr<Staging>> g_stage;

hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return hipSuccess;
  int dev = 0;
  hipError_t de = hipGetDevice(&dev);
  if (de != hipSuccess) return de;
  std::lock_guard<std::mutex> g(g_stage_mu);
  Staging* st = nullptr;
  for (auto& x : g_stage)
    if (x->device == dev && x->cap >= bytes && hipEventQuery(x->ev) == hipSuccess) {
      st = x.get();
      break;
    }
  if (!st) {
    auto n = std::make_unique<Staging>();
    n->device = dev;
    n->cap = std::max(bytes, (size_t)(64 << 10));
    hipError_t e = hipHostMalloc((void**)&n->p, n->cap, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    e = hipEventCreateWithFlags(&n->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipHostFree(n->p);
      return e;
    }
    st = n.get();
    g_stage.push_back(std::move(n));
  }
  std::memcpy(st->p, src, bytes);
  hipError_t e = hipMemcpyAsync(dst, st->p, bytes, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return e;
  return hipEventRecord(st->ev, stream);
}

// Launch one kernel over several instances: instances get consecutive
// workgroup ranges; one stream-ordered scratch block carries the instance
// table and every instance's refpt slot table.
// slots[i] is the refpt slot table of instance i; instances that share a
// table (same pointer) share its device copy.
int launch_instances(std::vector<Instance>& insts, const std::vector<const std::vector<int32_t>*>& slots,
                     const atx_launch_shape& shape, hipStream_t stream, bool strict, bool ext) {
  const int block = shape.block;
  const int per_wg = block * shape.ppt;  // particle slots per workgroup
  int nblocks = 0;
  for (auto& in : insts) {
    in.block_begin = nblocks;
    nblocks += (int)((in.npart + per_wg - 1) / per_wg);
  }
  if (nblocks == 0) return ATX_OK;
  size_t bytes = insts.size() * sizeof(Instance);
  std::vector<size_t> soff(insts.size());
  std::map<const std::vector<int32_t>*, size_t> placed;
  std::vector<std::pair<size_t, const std::vector<int32_t>*>> copies;
  for (size_t i = 0; i < insts.size(); i++) {
    auto f = placed.find(slots[i]);
    if (f != placed.end()) {
      soff[i] = f->second;
      continue;
    }
    soff[i] = bytes;
    placed[slots[i]] = bytes;
    copies.emplace_back(bytes, slots[i]);
    bytes += slots[i]->size() * sizeof(int32_t);
  }
  char* d_scratch = nullptr;
  HIPCHK(hipMallocAsync((void**)&d_scratch, bytes, stream));
  std::vector<char> host(bytes);
  for (size_t i = 0; i < insts.size(); i++) insts[i].slots = (const int32_t*)(d_scratch + soff[i]);
  for (auto& c : copies) std::memcpy(host.data() + c.first, c.second->data(), c.second->size() * sizeof(int32_t));
  std::memcpy(host.data(), insts.data(), insts.size() * sizeof(Instance));
  hipError_t ce = upload_async(d_scratch, host.data(), bytes, stream);
  hipError_t le = hipSuccess;
  if (ce == hipSuccess)
    le = track_launcher(strict, shape)((const Instance*)d_scratch, (int)insts.size(), nblocks, block,
                                       ext ? 1 : 0, xcd_order(insts.data(), insts.size()) ? 1 : 0,
                                       nullptr, stream);
  hipError_t fe = hipFreeAsync(d_scratch, stream);
  if (ce != hipSuccess) return fail(ATX_E_RUNTIME, std::string("upload: ") + hipGetErrorString(ce));
  if (le != hipSuccess) return fail(ATX_E_RUNTIME, std::string("kernel launch: ") + hipGetErrorString(le));
  if (fe != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipFreeAsync: ") + hipGetErrorString(fe));
  return ATX_OK;
}

// Graph executables of repacked launches: launched asynchronously, kept
// until the event recorded after them has completed, then destroyed.
struct PendingGraph {
  hipGraphExec_t exec;
  hipEvent_t done;
};
std::mutex g_graph_mu;
std::vector<PendingGraph> g_graphs;

void reap_graphs() {
  std::lock_guard<std::mutex> g(g_graph_mu);
  size_t k = 0;
  for (auto& pg : g_graphs) {
    if (hipEventQuery(pg.done) == hipSuccess) {
      (void)hipGraphExecDestroy(pg.exec);
      (void)hipEventDestroy(pg.done);
    } else {
      g_graphs[k++] = pg;
    }
  }
  g_graphs.resize(k);
}

// Turn-block boundaries of a repacked launch: 0, 1, 2, 4, ..., kMaxTurnBlock,
// then steps of kMaxTurnBlock, and nturns. The survivors are repacked and
// their workgroups placed anew at every boundary: a block keeps the
// placement its dispatch made, and with blocks doubling up to 512 turns a
// poor one in the latency-bound late blocks had cost up to a fifth of the
// block from one run to the next; short blocks also keep the survivors
// packed (each boundary costs about 0.04 ms of repack and map).
constexpr int kMaxTurnBlock = 16;
std::vector<int> repack_bounds(int nturns) {
  std::vector<int> b{0};
  for (int t = 1; t < nturns; t = t < kMaxTurnBlock ? 2 * t : t + kMaxTurnBlock) b.push_back(t);
  b.push_back(nturns);
  return b;
}

// Batched launch in turn blocks with the survivors of every instance packed
// into its leading slots between blocks (repack_kernel) and the workgroups
// holding them mapped to the front of the grid (live_map_kernel). All block
// tables go up in one copy; the track and repack launches are captured on a
// private stream into one graph, which runs on `stream`. Each particle's
// arithmetic is unchanged, so results equal those of launch_instances.
int launch_repacked(std::vector<Instance>& insts, const std::vector<const std::vector<int32_t>*>& slots,
                    const atx_launch_shape& shape, hipStream_t stream, bool strict, bool ext) {
  const int block = shape.block, ppt = shape.ppt;
  const int per_wg = block * ppt;  // particle slots per workgroup
  const int fine = 64 * ppt;       // ... when the survivors run one wave per workgroup
  reap_graphs();
  const size_t ninst = insts.size();
  const bool xcd = xcd_order(insts.data(), ninst);
  int nblocks = 0;
  int nblocks_fine = 0;  // grid of the later blocks: room for 64 particles per workgroup
  int maxt = 0;
  int64_t nidx = 0;
  for (auto& in : insts) {
    in.block_begin = nblocks;
    nblocks += (int)((in.npart + per_wg - 1) / per_wg);
    nblocks_fine += (int)((in.npart + fine - 1) / fine);
    maxt = std::max(maxt, (int)in.nturns);
    nidx += in.npart;
  }
  if (nblocks == 0 || maxt == 0) return ATX_OK;
  const std::vector<int> bounds = repack_bounds(maxt);
  const int nblk = (int)bounds.size() - 1;
  const size_t tab = ninst * sizeof(Instance);
  size_t bytes = (size_t)nblk * tab;
  std::vector<size_t> soff(ninst);
  std::map<const std::vector<int32_t>*, size_t> placed;
  std::vector<std::pair<size_t, const std::vector<int32_t>*>> copies;
  for (size_t i = 0; i < ninst; i++) {
    auto f = placed.find(slots[i]);
    if (f != placed.end()) {
      soff[i] = f->second;
      continue;
    }
    soff[i] = bytes;
    placed[slots[i]] = bytes;
    copies.emplace_back(bytes, slots[i]);
    bytes += slots[i]->size() * sizeof(int32_t);
  }
  const size_t upload = bytes;  // tables and slot tables; idx and nlive are written on the device
  bytes = (bytes + 255) & ~(size_t)255;
  const size_t idx_off = bytes;
  bytes += (size_t)nidx * sizeof(int32_t);
  const size_t nlive_off = bytes;
  bytes += ninst * sizeof(int32_t);
  const size_t live_off = bytes;  // live workgroup map of the current block
  bytes += (ninst + 2) * sizeof(int32_t);
  // survivors that fit in two waves per SIMD run one wave per workgroup
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 0;
  const int fine_limit = 2 * fine * 4 * ncu;
  char* d = nullptr;
  HIPCHK(hipMallocAsync((void**)&d, bytes, stream));
  std::vector<char> host(upload);
  for (auto& c : copies) std::memcpy(host.data() + c.first, c.second->data(), c.second->size() * sizeof(int32_t));
  for (int k = 0; k < nblk; k++) {
    int64_t ioff = 0;
    for (size_t i = 0; i < ninst; i++) {
      Instance in = insts[i];
      in.slots = (const int32_t*)(d + soff[i]);
      const int lo = bounds[k];
      const int hi = std::min(bounds[k + 1], (int)insts[i].nturns);
      in.nturns = std::max(0, hi - lo);
      in.turn0 = insts[i].turn0 + lo;
      // a partial first turn (start_elem) is the first block's only
      if (k > 0) in.start_elem = 0;
      if (k > 0) {
        in.losses = insts[i].losses ? 2 : 0;
        in.idx = (const int32_t*)(d + idx_off) + ioff;
        in.nlive = (const int32_t*)(d + nlive_off) + i;
      }
      ioff += insts[i].npart;
      std::memcpy(host.data() + (size_t)k * tab + i * sizeof(Instance), &in, sizeof(Instance));
    }
  }
  hipError_t ce = upload_async(d, host.data(), upload, stream);
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
  hipError_t ge = ce;
  if (ge == hipSuccess) ge = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (ge == hipSuccess) ge = hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed);
  if (ge == hipSuccess) {
    hipError_t le = hipSuccess;
    int32_t* live = (int32_t*)(d + live_off);
    for (int k = 0; k < nblk && le == hipSuccess; k++) {
      const Instance* t = (const Instance*)(d + (size_t)k * tab);
      // after the first block the survivors' workgroups lead the grid
      if (k > 0) le = atx_launch_live_map(t, (int)ninst, per_wg, fine, fine_limit, live, cap);
      if (le != hipSuccess) break;
      const int32_t* lb = k > 0 ? live : nullptr;
      const int grid = k > 0 ? nblocks_fine : nblocks;
      le = track_launcher(strict, shape)(t, (int)ninst, grid, block, ext ? 1 : 0, xcd ? 1 : 0, lb,
                                         cap);
      if (le == hipSuccess && k + 1 < nblk)
        le = atx_launch_repack((const Instance*)(d + (size_t)(k + 1) * tab), (int)ninst, k == 0 ? 1 : 0, cap);
    }
    ge = hipStreamEndCapture(cap, &graph);
    if (le != hipSuccess) ge = le;
  }
  if (ge == hipSuccess) ge = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (cap) (void)hipStreamDestroy(cap);
  if (ge == hipSuccess) ge = hipGraphLaunch(exec, stream);
  if (ge == hipSuccess) ge = hipEventCreateWithFlags(&done, hipEventDisableTiming);
  if (ge == hipSuccess) ge = hipEventRecord(done, stream);
  if (ge == hipSuccess) {
    std::lock_guard<std::mutex> g(g_graph_mu);
    g_graphs.push_back({exec, done});
  } else {
    // nothing of the graph ran, or it has been queued before the event failed
    if (done) (void)hipEventDestroy(done);
    if (exec) {
      (void)hipStreamSynchronize(stream);
      (void)hipGraphExecDestroy(exec);
    }
  }
  hipError_t fe = hipFreeAsync(d, stream);
  if (ce != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipMemcpyAsync: ") + hipGetErrorString(ce));
  if (ge != hipSuccess) return fail(ATX_E_RUNTIME, std::string("repacked launch: ") + hipGetErrorString(ge));
  if (fe != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipFreeAsync: ") + hipGetErrorString(fe));
  return ATX_OK;
}

// The instance of one atx_track call and its refpt slot table (first turn,
// then later turns), with the argument checks that need them.
int single_instance(atx_lattice* lat, double* d_r, int64_t pstride, int64_t cstride, uint32_t np,
                    int nturns, const uint32_t* refpts, uint32_t nref, int turn0, int start, int end,
                    double* d_rout, uint8_t* d_lf, uint32_t* d_lt, uint32_t* d_le, double* d_lc,
                    int losses, bool elempass, int32_t* slots, Instance& inst) {
  const int nelem = (int)lat->elems.size();
  const int n_first = turn_slots(refpts, (int)nref, start, end, nelem, slots);
  const int n_std = turn_slots(refpts, (int)nref, 0, end, nelem, slots + nelem + 1);
  if ((n_first > 0 || n_std > 0) && !d_rout) return fail(ATX_E_VALUE, "r_out is NULL but refpts request copies");
  if (losses && (!d_lf || !d_lt || !d_le || !d_lc)) return fail(ATX_E_VALUE, "loss arrays are NULL");
  std::memset(&inst, 0, sizeof(inst));
  inst.elems = lat->d_elems;
  inst.pool = lat->d_pool;
  inst.r = d_r;
  inst.rout = d_rout;
  inst.lost_flag = d_lf;
  inst.lost_turn = d_lt;
  inst.lost_elem = d_le;
  inst.lost_coord = d_lc;
  inst.pstride = pstride;
  inst.cstride = cstride;
  inst.nelem = nelem;
  inst.npart = (int32_t)np;
  inst.nturns = nturns;
  inst.turn0 = turn0;
  inst.start_elem = start;
  inst.end_elem = end;
  inst.n_first = n_first;
  inst.n_std = n_std;
  inst.block_begin = 0;
  inst.losses = losses ? 1 : 0;
  inst.elempass = elempass ? 1 : 0;
  return ATX_OK;
}

// Device-side core of atx_track: all pointers are device pointers.
int track_device(atx_lattice* lat, double* d_r, int64_t pstride, int64_t cstride, uint32_t np,
                 int nturns, const uint32_t* refpts, uint32_t nref, int turn0, int start, int end,
                 double* d_rout, uint8_t* d_lf, uint32_t* d_lt, uint32_t* d_le, double* d_lc,
                 int losses, bool elempass, hipStream_t stream, bool strict, bool repack = false) {
  const int nelem = (int)lat->elems.size();
  std::vector<std::vector<int32_t>> slots(1, std::vector<int32_t>(2 * (size_t)(nelem + 1)));
  std::vector<Instance> insts(1);
  int rc = single_instance(lat, d_r, pstride, cstride, np, nturns, refpts, nref, turn0, start, end,
                           d_rout, d_lf, d_lt, d_le, d_lc, losses, elempass, slots[0].data(), insts[0]);
  if (rc) return rc;
  if (np == 0 || nturns <= 0) return ATX_OK;
  atx_launch_shape shape;
  rc = resolve_shape(&lat->shape, shape, strict);
  if (rc) return rc;
  // Without a shape set on the lattice, beams of two or more 512-lane
  // workgroups per CU run in 512-lane workgroups: survivors are packed
  // within a workgroup, and larger ones pack more of them (2 % faster on the
  // large ring at 262,144 particles than 256 lanes).
  int dev = 0, ncu = 0;
  if (lat->shape.block == 0 && hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      (int64_t)np >= (int64_t)2 * 512 * shape.ppt * ncu)
    shape.block = 512;
  // survivors repacked between turn blocks: only without refpts (lost
  // particles still fill their copies, at.c:622-626)
  if (repack && nref == 0 && nturns > 1)
    return launch_repacked(insts, {&slots[0]}, shape, stream, strict, lat->has_ext);
  return launch_instances(insts, {&slots[0]}, shape, stream, strict, lat->has_ext);
}

int check_partial(int nelem, int nturns, int& start, int& end) {
  // at.c:514-536
  if (end == 0) end = nelem;
  if (nelem != 0) {
    if (start < 0 || start >= nelem)
      return fail(ATX_E_VALUE, "start_elem out of range, " + std::to_string(start) + " not in [0.." +
                                   std::to_string(nelem - 1) + "]");
    if (end < 0 || end > nelem)
      return fail(ATX_E_VALUE, "end_elem out of range, " + std::to_string(end) + " not in [0.." +
                                   std::to_string(nelem) + "]");
    if (start > end) return fail(ATX_E_VALUE, "end_elem must be greater that start_elem");
    if (end != nelem && nturns != 1)
      return fail(ATX_E_VALUE, "partial turn tracking not ending at the end of the ring requires nturns=1");
  } else {
    start = 0;
    end = 0;
  }
  return ATX_OK;
}

}  // namespace
}  // namespace atx

using namespace atx;

// --- Instance table of atx_track_batch_noalloc --------------------------------
// A file-scope device array, one per GPU, holding the instance table of the
// last zero-allocation launch on that GPU. In host code `scratch` only names
// the symbol: its device address is taken once per device with
// hipGetSymbolAddress and used from then on.
constexpr int kNoAllocItems = 64;
static __device__ __align__(256) Instance scratch[kNoAllocItems];

namespace atx {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
Instance* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (first call on a device:
// hipGetSymbolAddress, with that device current).
int scratch_address(Instance** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ATX_E_RUNTIME, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ATX_E_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<Instance*>(p);
  }
  *out = g_scratch_ptr[dev];
  return ATX_OK;
}

// Instances travel to the device as kernel arguments, 16 per launch.
constexpr int kChunk = 16;
struct InstanceChunk {
  Instance v[kChunk];
};

__global__ void __launch_bounds__(64) put_instances(Instance* dst, InstanceChunk c, int n) {
  const int i = (int)threadIdx.x;
  if (i < n) dst[i] = c.v[i];
}

}  // namespace
}  // namespace atx

extern "C" {

const char* atx_last_error(void) { return g_err.c_str(); }

int atx_lattice_create(atx_lattice** out) {
  if (!out) return fail(ATX_E_VALUE, "NULL handle pointer");
  *out = new atx_lattice();
  return ATX_OK;
}

int atx_lattice_destroy(atx_lattice* lat) {
  if (!lat) return ATX_OK;
  if (lat->d_elems) (void)hipFree(lat->d_elems);
  if (lat->d_pool) (void)hipFree(lat->d_pool);
  delete lat;
  return ATX_OK;
}

int atx_lattice_add_element(atx_lattice* lat, const char* pass_method, const char* fam_name) {
  if (!lat || !pass_method) return fail(ATX_E_VALUE, "NULL argument");
  for (const auto& p : kPasses) {
    if (std::strcmp(p.name, pass_method) == 0) {
      std::lock_guard<std::mutex> g(lat->mu);
      HostElem e;
      e.pass = pass_method;
      e.fam = fam_name ? fam_name : "";
      e.kind = p.kind;
      lat->elems.push_back(std::move(e));
      lat->dev_valid = false;
      return (int)ATX_OK;
    }
  }
  return fail(ATX_E_RUNTIME, std::string("Not implemented PassMethod: ") + pass_method +
                                 " in element #" + std::to_string(lat->elems.size()) + " (" +
                                 (fam_name ? fam_name : "") + ")");
}

static int set_attr(atx_lattice* lat, const char* name, const double* v, int64_t n, bool is_int) {
  if (!lat || !name) return fail(ATX_E_VALUE, "NULL argument");
  if (lat->elems.empty()) return fail(ATX_E_VALUE, "no element to set an attribute on");
  if (n < 0 || (n > 0 && !v)) return fail(ATX_E_VALUE, "bad array");
  std::lock_guard<std::mutex> g(lat->mu);
  Attr a;
  a.v.assign(v, v + n);
  a.is_int = is_int;
  lat->elems.back().attrs[name] = std::move(a);
  lat->dev_valid = false;
  return ATX_OK;
}

int atx_element_set_double(atx_lattice* lat, const char* name, double value) {
  return set_attr(lat, name, &value, 1, false);
}

int atx_element_set_int(atx_lattice* lat, const char* name, int64_t value) {
  double v = (double)value;
  return set_attr(lat, name, &v, 1, true);
}

int atx_element_set_array(atx_lattice* lat, const char* name, const double* data, int64_t n) {
  return set_attr(lat, name, data, n, false);
}

int atx_lattice_num_elements(const atx_lattice* lat) { return lat ? (int)lat->elems.size() : 0; }

int atx_lattice_set_integrator(atx_lattice* lat, int32_t type) {
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  Scheme s;
  if (!integrator_scheme(type, &s))
    return fail(ATX_E_RUNTIME, "SymplecticIntegrator, type " + std::to_string(type) + " not supported");
  std::lock_guard<std::mutex> g(lat->mu);
  if (lat->integrator != type) {
    lat->integrator = type;
    lat->dev_valid = false;  // recompiled, as atgpu recreates its lattice (PyInterface.cpp:259-266)
  }
  return ATX_OK;
}

int atx_lattice_check(atx_lattice* lat, double energy, double rest_energy) {
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  Compiled c;
  return compile_lattice(lat, energy, rest_energy, false, c);
}

int atx_lattice_turn_counter(const atx_lattice* lat, int64_t* turn) {
  if (!lat || !turn) return fail(ATX_E_VALUE, "NULL argument");
  *turn = lat->turn_counter;
  return ATX_OK;
}

int atx_track(atx_lattice* lat, double* r_in, uint32_t np, int32_t nturns, const uint32_t* refpts,
              uint32_t nref, int32_t turn, double energy, double rest_energy, double charge,
              int32_t losses, int32_t start_elem, int32_t end_elem, double* r_out,
              uint8_t* lost_flag, uint32_t* lost_turn, uint32_t* lost_elem, double* lost_coord,
              void* stream_, uint32_t flags) {
  (void)charge;
  g_err.clear();
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  if (np > 0 && !r_in) return fail(ATX_E_VALUE, "r_in is NULL");
  if (nturns < 0) return fail(ATX_E_VALUE, "nturns must be >= 0");
  if (nref > 0 && !refpts) return fail(ATX_E_VALUE, "refpts is NULL");
  for (uint32_t i = 1; i < nref; i++)
    if (refpts[i] < refpts[i - 1]) return fail(ATX_E_VALUE, "refpts must be sorted");
  std::lock_guard<std::mutex> g(lat->mu);
  const int nelem = (int)lat->elems.size();
  int start = start_elem, end = end_elem;
  int rc = check_partial(nelem, nturns, start, end);
  if (rc) return rc;
  rc = ensure_device_lattice(lat, energy, rest_energy, false);
  if (rc) return rc;
  const int turn0 = (flags & ATX_KEEP_COUNTER) ? (int)lat->turn_counter : turn;
  hipStream_t stream = (hipStream_t)stream_;
  const bool strict = (flags & ATX_STRICT) != 0;

  // the (6,N,R,T) copy count the reference produces
  std::vector<int32_t> tmp(nelem + 1);
  const int n_first = nturns > 0 ? turn_slots(refpts, (int)nref, start, end, nelem, tmp.data()) : 0;
  const int n_std = turn_slots(refpts, (int)nref, 0, end, nelem, tmp.data());
  (void)n_first;
  (void)n_std;

  if (flags & ATX_DEVICE_POINTERS) {
    const bool soa = (flags & ATX_SOA) != 0;
    rc = track_device(lat, r_in, soa ? 1 : 6, soa ? (int64_t)np : 1, np, nturns, refpts, nref, turn0,
                      start, end, r_out, lost_flag, lost_turn, lost_elem, lost_coord, losses, false,
                      stream, strict, (flags & ATX_REPACK) != 0);
    if (rc) return rc;
  } else {
    // host buffers: stage through device memory
    const size_t nr = 6 * (size_t)np;
    const size_t nout = nr * (size_t)nref * (size_t)nturns;
    double *d_r = nullptr, *d_out = nullptr, *d_lc = nullptr;
    uint8_t* d_lf = nullptr;
    uint32_t *d_lt = nullptr, *d_le = nullptr;
    auto cleanup = [&]() {
      if (d_r) (void)hipFree(d_r);
      if (d_out) (void)hipFree(d_out);
      if (d_lc) (void)hipFree(d_lc);
      if (d_lf) (void)hipFree(d_lf);
      if (d_lt) (void)hipFree(d_lt);
      if (d_le) (void)hipFree(d_le);
    };
#define HIPCHK2(x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) {                                                                 \
      cleanup();                                                                            \
      return fail(ATX_E_RUNTIME, std::string(#x) + ": " + hipGetErrorString(e_));           \
    }                                                                                       \
  } while (0)
    if (np > 0) {
      HIPCHK2(hipMalloc(&d_r, nr * sizeof(double)));
      HIPCHK2(hipMemcpy(d_r, r_in, nr * sizeof(double), hipMemcpyHostToDevice));
    }
    if (nout > 0 && r_out) {
      HIPCHK2(hipMalloc(&d_out, nout * sizeof(double)));
      // slots the reference never writes keep the caller's content
      HIPCHK2(hipMemcpy(d_out, r_out, nout * sizeof(double), hipMemcpyHostToDevice));
    }
    if (losses && np > 0) {
      HIPCHK2(hipMalloc(&d_lc, nr * sizeof(double)));
      HIPCHK2(hipMalloc(&d_lf, np));
      HIPCHK2(hipMalloc(&d_lt, np * sizeof(uint32_t)));
      HIPCHK2(hipMalloc(&d_le, np * sizeof(uint32_t)));
    }
    rc = track_device(lat, d_r, 6, 1, np, nturns, refpts, nref, turn0, start, end, d_out, d_lf, d_lt,
                      d_le, d_lc, losses, false, stream, strict, (flags & ATX_REPACK) != 0);
    if (rc) {
      cleanup();
      return rc;
    }
    HIPCHK2(hipStreamSynchronize(stream));
    if (np > 0) HIPCHK2(hipMemcpy(r_in, d_r, nr * sizeof(double), hipMemcpyDeviceToHost));
    if (d_out) HIPCHK2(hipMemcpy(r_out, d_out, nout * sizeof(double), hipMemcpyDeviceToHost));
    if (losses && np > 0) {
      HIPCHK2(hipMemcpy(lost_coord, d_lc, nr * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK2(hipMemcpy(lost_flag, d_lf, np, hipMemcpyDeviceToHost));
      HIPCHK2(hipMemcpy(lost_turn, d_lt, np * sizeof(uint32_t), hipMemcpyDeviceToHost));
      HIPCHK2(hipMemcpy(lost_elem, d_le, np * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    cleanup();
#undef HIPCHK2
  }
  lat->turn_counter = (int64_t)turn0 + nturns;  // at.c:652,656
  return ATX_OK;
}

int atx_shutdown(void) {
  g_err.clear();
  hipError_t first = hipSuccess;
  auto keep = [&](hipError_t e) {
    if (first == hipSuccess && e != hipSuccess) first = e;
  };
  {
    std::lock_guard<std::mutex> g(g_graph_mu);
    for (auto& pg : g_graphs) {
      keep(hipEventSynchronize(pg.done));
      keep(hipGraphExecDestroy(pg.exec));
      keep(hipEventDestroy(pg.done));
    }
    g_graphs.clear();
  }
  {
    std::lock_guard<std::mutex> g(g_stage_mu);
    for (auto& st : g_stage) {
      keep(hipEventSynchronize(st->ev));
      keep(hipEventDestroy(st->ev));
      keep(hipHostFree(st->p));
    }
    g_stage.clear();
  }
  if (first != hipSuccess) return fail(ATX_E_RUNTIME, std::string("atx_shutdown: ") + hipGetErrorString(first));
  return ATX_OK;
}

int atx_track_batch(const atx_batch_item* items, int32_t nitems, int32_t nturns, int32_t turn,
                    double energy, double rest_energy, int32_t losses,
                    const atx_launch_shape* shape_in, void* stream_, uint32_t flags) {
  g_err.clear();
  if (nitems < 0 || (nitems > 0 && !items)) return fail(ATX_E_VALUE, "bad item list");
  if (nturns < 0) return fail(ATX_E_VALUE, "nturns must be >= 0");
  const bool strict = (flags & ATX_STRICT) != 0;
  atx_launch_shape shape;
  if (int rc = resolve_shape(shape_in, shape, strict)) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  const bool soa = (flags & ATX_SOA) != 0;
  for (int i = 0; i < nitems; i++)
    if (items[i].nturns >= 0 && items[i].turn < 0)
      return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": turn must be >= 0");
  std::vector<Instance> insts;
  bool ext = false;
  std::map<int, std::vector<int32_t>> empty_slots;  // refpts = [] table per ring length
  std::vector<const std::vector<int32_t>*> slots;
  for (int i = 0; i < nitems; i++) {
    const atx_batch_item& it = items[i];
    if (!it.lattice) return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": NULL lattice");
    if (it.num_particles > 0 && !it.r) return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": NULL r");
    if (losses && it.num_particles > 0 && (!it.lost_flag || !it.lost_turn || !it.lost_elem || !it.lost_coord))
      return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": loss arrays are NULL");
  }
  // Each distinct lattice is locked once, in address order: two calls that
  // list the same lattices in different orders take the locks in the same
  // order and cannot deadlock.
  std::vector<atx_lattice*> distinct;
  for (int i = 0; i < nitems; i++) distinct.push_back(items[i].lattice);
  std::sort(distinct.begin(), distinct.end(), std::less<atx_lattice*>());
  distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
  std::vector<std::unique_lock<std::mutex>> locks;
  locks.reserve(distinct.size());
  for (atx_lattice* l : distinct) locks.emplace_back(l->mu);
  for (int i = 0; i < nitems; i++) {
    const atx_batch_item& it = items[i];
    int rc = ensure_device_lattice(it.lattice, energy, rest_energy, false);
    if (rc) return fail(rc, "item " + std::to_string(i) + ": " + g_err);
    if (it.num_particles == 0) continue;
    ext |= it.lattice->has_ext;
    const int nelem = (int)it.lattice->elems.size();
    Instance in;
    std::memset(&in, 0, sizeof(in));
    in.elems = it.lattice->d_elems;
    in.pool = it.lattice->d_pool;
    in.r = it.r;
    in.lost_flag = it.lost_flag;
    in.lost_turn = it.lost_turn;
    in.lost_elem = it.lost_elem;
    in.lost_coord = it.lost_coord;
    in.pstride = soa ? 1 : 6;
    in.cstride = soa ? (int64_t)it.num_particles : 1;
    in.nelem = nelem;
    in.npart = (int32_t)it.num_particles;
    in.nturns = it.nturns >= 0 ? it.nturns : nturns;
    in.turn0 = it.nturns >= 0 ? it.turn : turn;
    in.start_elem = 0;
    in.end_elem = nelem;
    in.losses = losses ? ((flags & ATX_CONTINUE_LOSSES) ? 2 : 1) : 0;
    if (in.nturns == 0) continue;
    insts.push_back(in);
    auto ins = empty_slots.emplace(nelem, std::vector<int32_t>(2 * (size_t)(nelem + 1), -1));
    slots.push_back(&ins.first->second);
  }
  if (insts.empty()) return ATX_OK;
  // Longest work per workgroup first: the hardware dispatches workgroups in
  // index order, so long instances start early and short ones fill the tail.
  std::vector<size_t> order(insts.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return (int64_t)insts[a].nturns * insts[a].nelem > (int64_t)insts[b].nturns * insts[b].nelem;
  });
  std::vector<Instance> sorted_insts;
  std::vector<const std::vector<int32_t>*> sorted_slots;
  sorted_insts.reserve(insts.size());
  sorted_slots.reserve(insts.size());
  for (size_t i : order) {
    sorted_insts.push_back(insts[i]);
    sorted_slots.push_back(slots[i]);
  }
  if (flags & ATX_REPACK)
    return launch_repacked(sorted_insts, sorted_slots, shape, stream, strict, ext);
  return launch_instances(sorted_insts, sorted_slots, shape, stream, strict, ext);
}

int atx_track_batch_noalloc(const atx_batch_item* items, int32_t nitems, int32_t nturns,
                            int32_t turn, double energy, double rest_energy, int32_t losses,
                            const atx_launch_shape* shape_in, void* stream_, uint32_t flags) {
  g_err.clear();
  if (nitems < 0 || (nitems > 0 && !items)) return fail(ATX_E_VALUE, "bad item list");
  if (nitems > kNoAllocItems)
    return fail(ATX_E_VALUE, "atx_track_batch_noalloc takes at most 64 items");
  if (nturns < 0) return fail(ATX_E_VALUE, "nturns must be >= 0");
  if (flags & ATX_REPACK)
    return fail(ATX_E_VALUE, "ATX_REPACK needs working memory: not available without allocation");
  const bool strict = (flags & ATX_STRICT) != 0;
  atx_launch_shape shape;
  if (int rc = resolve_shape(shape_in, shape, strict)) return rc;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  Instance* d_tab = nullptr;
  if (int rc = scratch_address(&d_tab)) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  const bool soa = (flags & ATX_SOA) != 0;
  for (int i = 0; i < nitems; i++) {
    const atx_batch_item& it = items[i];
    if (!it.lattice) return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": NULL lattice");
    if (it.num_particles > 0 && !it.r) return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": NULL r");
    if (losses && it.num_particles > 0 && (!it.lost_flag || !it.lost_turn || !it.lost_elem || !it.lost_coord))
      return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": loss arrays are NULL");
    if (it.nturns >= 0 && it.turn < 0)
      return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": turn must be >= 0");
  }
  // each distinct lattice locked once, in address order (as atx_track_batch)
  atx_lattice* distinct[kNoAllocItems];
  int nd = 0;
  for (int i = 0; i < nitems; i++) distinct[nd++] = items[i].lattice;
  std::sort(distinct, distinct + nd, std::less<atx_lattice*>());
  nd = (int)(std::unique(distinct, distinct + nd) - distinct);
  struct Locks {
    atx_lattice** l;
    int n = 0;
    ~Locks() {
      for (int i = n - 1; i >= 0; i--) l[i]->mu.unlock();
    }
  } held{distinct};
  for (int i = 0; i < nd; i++) {
    distinct[i]->mu.lock();
    held.n = i + 1;
  }
  Instance tab[kNoAllocItems];
  int64_t work[kNoAllocItems];
  int n = 0;
  bool ext = false;
  for (int i = 0; i < nitems; i++) {
    const atx_batch_item& it = items[i];
    atx_lattice* lat = it.lattice;
    // compiling or uploading a lattice allocates: it must already be on
    // this device for this energy
    if (!(lat->dev_valid && lat->dev_energy == energy && lat->dev_rest == rest_energy &&
          !lat->dev_elempass && lat->dev_device == dev))
      return fail(ATX_E_VALUE, "item " + std::to_string(i) +
                                   ": lattice not on this device for this energy (track it once "
                                   "with atx_track or atx_track_batch first)");
    const int nt = it.nturns >= 0 ? it.nturns : nturns;
    if (it.num_particles == 0 || nt == 0) continue;
    ext |= lat->has_ext;
    Instance& in = tab[n];
    std::memset(&in, 0, sizeof(in));
    const int nelem = (int)lat->elems.size();
    in.elems = lat->d_elems;
    in.pool = lat->d_pool;
    in.slots = nullptr;  // no refpts: the kernel reads no slot table
    in.r = it.r;
    in.lost_flag = it.lost_flag;
    in.lost_turn = it.lost_turn;
    in.lost_elem = it.lost_elem;
    in.lost_coord = it.lost_coord;
    in.pstride = soa ? 1 : 6;
    in.cstride = soa ? (int64_t)it.num_particles : 1;
    in.nelem = nelem;
    in.npart = (int32_t)it.num_particles;
    in.nturns = nt;
    in.turn0 = it.nturns >= 0 ? it.turn : turn;
    in.start_elem = 0;
    in.end_elem = nelem;
    in.losses = losses ? ((flags & ATX_CONTINUE_LOSSES) ? 2 : 1) : 0;
    work[n] = (int64_t)nt * nelem;
    n++;
  }
  if (n == 0) return ATX_OK;
  // longest work first (insertion sort: stable, in place)
  for (int i = 1; i < n; i++) {
    const Instance v = tab[i];
    const int64_t w = work[i];
    int j = i - 1;
    for (; j >= 0 && work[j] < w; j--) {
      tab[j + 1] = tab[j];
      work[j + 1] = work[j];
    }
    tab[j + 1] = v;
    work[j + 1] = w;
  }
  const int per_wg = shape.block * shape.ppt;
  int nblocks = 0;
  for (int i = 0; i < n; i++) {
    tab[i].block_begin = nblocks;
    nblocks += (int)((tab[i].npart + per_wg - 1) / per_wg);
  }
  for (int c = 0; c < n; c += kChunk) {
    InstanceChunk chunk;
    const int m = std::min(kChunk, n - c);
    for (int k = 0; k < m; k++) chunk.v[k] = tab[c + k];
    hipLaunchKernelGGL(put_instances, dim3(1), dim3(64), 0, stream, d_tab + c, chunk, m);
    HIPCHK(hipGetLastError());
  }
  const hipError_t le = track_launcher(strict, shape)(d_tab, n, nblocks, shape.block, ext ? 1 : 0,
                                                      xcd_order(tab, (size_t)n) ? 1 : 0, nullptr,
                                                      stream);
  if (le != hipSuccess) return fail(ATX_E_RUNTIME, std::string("kernel launch: ") + hipGetErrorString(le));
  return ATX_OK;
}

int atx_lattice_set_launch_shape(atx_lattice* lat, const atx_launch_shape* shape) {
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  atx_launch_shape s;
  if (int rc = resolve_shape(shape, s, false)) return rc;
  std::lock_guard<std::mutex> g(lat->mu);
  lat->shape = shape ? *shape : atx_launch_shape{0, 0, 0, 0};
  return ATX_OK;
}

int atx_element_set_array_at(atx_lattice* lat, int32_t index, const char* name, const double* data,
                             int64_t n) {
  if (!lat || !name) return fail(ATX_E_VALUE, "NULL argument");
  if (n < 0 || (n > 0 && !data)) return fail(ATX_E_VALUE, "bad array");
  std::lock_guard<std::mutex> g(lat->mu);
  if (index < 0 || index >= (int32_t)lat->elems.size()) return fail(ATX_E_VALUE, "element index out of range");
  Attr a;
  a.v.assign(data, data + n);
  lat->elems[index].attrs[name] = std::move(a);
  lat->dev_valid = false;
  return ATX_OK;
}

// Shared by atx_adjoint and atx_adjoint_batch: per-element adjoint table,
// one upload of that table and the item table, one launch.
static int adjoint_launch(atx_lattice* lat, const std::vector<AdjItem>& items, double energy,
                          double rest_energy, const int32_t* grad_slot, const int32_t* grad_coeff,
                          bool have_grad, int32_t num_params, void* stream_) {
  std::lock_guard<std::mutex> g(lat->mu);
  const int nelem = (int)lat->elems.size();
  int rc = ensure_device_lattice(lat, energy, rest_energy, false);
  if (rc) return rc;
  std::vector<AdjElem> ae(nelem > 0 ? nelem : 1);
  for (int e = 0; e < nelem; e++) {
    if (!lat->h_adjerr[e].empty())
      return fail(ATX_E_RUNTIME, "adjoint not available for element #" + std::to_string(e) + " " +
                                     lat->elems[e].fam + " (" + lat->elems[e].pass + "): " + lat->h_adjerr[e]);
    const int slot = grad_slot ? grad_slot[e] : -1;
    const int coeff = grad_coeff ? grad_coeff[e] : 0;
    if (slot >= 0) {
      const int k = lat->h_kind[e];
      if (k != K_STR && k != K_BND && k != K_THIN)
        return fail(ATX_E_VALUE, "element #" + std::to_string(e) + " has no PolynomB to differentiate");
      if (slot >= num_params || !have_grad) return fail(ATX_E_VALUE, "gradient slot out of range");
      if (coeff < 0 || coeff > lat->h_order[e])
        return fail(ATX_E_VALUE, "element #" + std::to_string(e) + ": PolynomB index above MaxOrder");
    }
    ae[e].slot = slot;
    ae[e].coeff = coeff;
    ae[e].r1inv = lat->h_r1inv[e];
    ae[e].r2inv = lat->h_r2inv[e];
  }
  std::vector<AdjItem> run;
  int32_t max_np = 0;
  for (const AdjItem& it : items)
    if (it.npart > 0 && it.nturns > 0) {
      run.push_back(it);
      max_np = std::max(max_np, it.npart);
    }
  if (run.empty()) return ATX_OK;
  // one buffer: the element table, then the items (both 8-byte aligned)
  const size_t ae_bytes = ae.size() * sizeof(AdjElem);
  const size_t ae_pad = (ae_bytes + 255) & ~(size_t)255;
  std::vector<unsigned char> host(ae_pad + run.size() * sizeof(AdjItem));
  std::memcpy(host.data(), ae.data(), ae_bytes);
  std::memcpy(host.data() + ae_pad, run.data(), run.size() * sizeof(AdjItem));
  hipStream_t stream = (hipStream_t)stream_;
  unsigned char* d_buf = nullptr;
  HIPCHK(hipMallocAsync((void**)&d_buf, host.size(), stream));
  hipError_t ce = upload_async(d_buf, host.data(), host.size(), stream);
  hipError_t le = hipSuccess;
  if (ce == hipSuccess) {
    AdjArgs a;
    a.elems = lat->d_elems;
    a.pool = lat->d_pool;
    a.adj = (const AdjElem*)d_buf;
    a.items = (const AdjItem*)(d_buf + ae_pad);
    a.nelem = nelem;
    a.nitems = (int32_t)run.size();
    le = atx_launch_adjoint(&a, max_np, stream);
  }
  hipError_t fe = hipFreeAsync(d_buf, stream);
  if (ce != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipMemcpyAsync: ") + hipGetErrorString(ce));
  if (le != hipSuccess) return fail(ATX_E_RUNTIME, std::string("adjoint launch: ") + hipGetErrorString(le));
  if (fe != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipFreeAsync: ") + hipGetErrorString(fe));
  return ATX_OK;
}

int atx_adjoint(atx_lattice* lat, const double* states, int32_t nturns, int32_t turn, uint32_t np,
                double energy, double rest_energy, double* lam, const int32_t* grad_slot,
                const int32_t* grad_coeff, double* grad, int32_t num_params, void* stream_) {
  g_err.clear();
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  if (nturns < 0) return fail(ATX_E_VALUE, "nturns must be >= 0");
  if (np > 0 && (!states || !lam)) return fail(ATX_E_VALUE, "states/lam are NULL");
  if (np > (uint32_t)INT32_MAX) return fail(ATX_E_VALUE, "num_particles too large");
  AdjItem it{};
  it.states = states;
  it.lam = lam;
  it.grad = grad;
  it.npart = (int32_t)np;
  it.nturns = nturns;
  it.turn0 = turn;
  return adjoint_launch(lat, {it}, energy, rest_energy, grad_slot, grad_coeff, grad != nullptr,
                        num_params, stream_);
}

int atx_adjoint_batch(atx_lattice* lat, const atx_adjoint_item* items, int32_t nitems, double energy,
                      double rest_energy, const int32_t* grad_slot, const int32_t* grad_coeff,
                      double* grad, int32_t num_params, void* stream_) {
  g_err.clear();
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  if (nitems < 0 || (nitems > 0 && !items)) return fail(ATX_E_VALUE, "bad item list");
  if (nitems > 65535) return fail(ATX_E_VALUE, "more than 65535 items");
  std::vector<AdjItem> v;
  v.reserve((size_t)nitems);
  bool accumulate = false;
  for (int32_t i = 0; i < nitems; i++) {
    const atx_adjoint_item& s = items[i];
    if (s.nturns < 0) return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": nturns must be >= 0");
    if (s.num_particles > (uint32_t)INT32_MAX)
      return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": num_particles too large");
    if (s.num_particles > 0 && (!s.states || !s.lam))
      return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": states/lam are NULL");
    if (s.accumulate && !grad) return fail(ATX_E_VALUE, "item " + std::to_string(i) + ": grad is NULL");
    AdjItem it{};
    it.states = s.states;
    it.lam = s.lam;
    it.grad = s.accumulate ? grad : nullptr;
    accumulate = accumulate || s.accumulate;
    it.npart = (int32_t)s.num_particles;
    it.nturns = s.nturns;
    it.turn0 = s.turn;
    v.push_back(it);
  }
  // gradient slots need a gradient array only when some item accumulates
  return adjoint_launch(lat, v, energy, rest_energy, grad_slot, grad_coeff,
                        grad != nullptr || !accumulate, num_params, stream_);
}

int atx_diffusion_matrices(atx_lattice* lat, const double* orbits, double* bdiff, double energy,
                           double rest_energy, void* stream_, uint32_t flags) {
  g_err.clear();
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  std::lock_guard<std::mutex> g(lat->mu);
  const int nelem = (int)lat->elems.size();
  if (nelem == 0) return ATX_OK;
  if (!orbits || !bdiff) return fail(ATX_E_VALUE, "orbits/bdiff are NULL");
  int rc = ensure_device_lattice(lat, energy, rest_energy, false);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  const bool dev = (flags & ATX_DEVICE_POINTERS) != 0;
  const size_t no = 6 * (size_t)nelem, nb = 36 * (size_t)nelem;
  double* d_o = nullptr;
  double* d_b = nullptr;
  if (dev) {
    d_o = const_cast<double*>(orbits);
    d_b = bdiff;
  } else {
    HIPCHK(hipMallocAsync((void**)&d_o, (no + nb) * sizeof(double), stream));
    d_b = d_o + no;
    hipError_t ce = hipMemcpyAsync(d_o, orbits, no * sizeof(double), hipMemcpyHostToDevice, stream);
    if (ce != hipSuccess) {
      (void)hipFreeAsync(d_o, stream);
      return fail(ATX_E_RUNTIME, std::string("hipMemcpyAsync: ") + hipGetErrorString(ce));
    }
  }
  DiffArgs a;
  a.elems = lat->d_elems;
  a.pool = lat->d_pool;
  a.orbits = d_o;
  a.bdiff = d_b;
  a.nelem = nelem;
  a.pad = 0;
  hipError_t le = atx_launch_diffusion(&a, stream);
  if (dev) {
    if (le != hipSuccess) return fail(ATX_E_RUNTIME, std::string("diffusion launch: ") + hipGetErrorString(le));
    return ATX_OK;
  }
  hipError_t ce = hipSuccess;
  if (le == hipSuccess) ce = hipMemcpyAsync(bdiff, d_b, nb * sizeof(double), hipMemcpyDeviceToHost, stream);
  hipError_t fe = hipFreeAsync(d_o, stream);
  hipError_t se = hipStreamSynchronize(stream);
  if (le != hipSuccess) return fail(ATX_E_RUNTIME, std::string("diffusion launch: ") + hipGetErrorString(le));
  if (ce != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipMemcpyAsync: ") + hipGetErrorString(ce));
  if (fe != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipFreeAsync: ") + hipGetErrorString(fe));
  if (se != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
  return ATX_OK;
}

int atx_element_matrices(atx_lattice* lat, const double* orbits, double step, double* m66,
                         double energy, double rest_energy, void* stream_, uint32_t flags) {
  g_err.clear();
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  std::lock_guard<std::mutex> g(lat->mu);
  const int nelem = (int)lat->elems.size();
  if (nelem == 0) return ATX_OK;
  if (!orbits || !m66) return fail(ATX_E_VALUE, "orbits/m66 are NULL");
  if (!(step != 0.0) || !std::isfinite(step)) return fail(ATX_E_VALUE, "step must be finite and nonzero");
  int rc = ensure_device_lattice(lat, energy, rest_energy, true);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  const bool dev = (flags & ATX_DEVICE_POINTERS) != 0;
  const bool strict = (flags & ATX_STRICT) != 0;
  // probes (6, 12 nelem), then (host arrays) the orbits and the matrices
  const size_t nr = 72 * (size_t)nelem, no = 6 * (size_t)nelem, nm = 36 * (size_t)nelem;
  double* d_r = nullptr;
  HIPCHK(hipMallocAsync((void**)&d_r, (nr + (dev ? 0 : no + nm)) * sizeof(double), stream));
  const double* d_o = dev ? orbits : d_r + nr;
  double* d_m = dev ? m66 : d_r + nr + no;
  // One instance per element: its 12 probes through that element alone
  // (at_elempass: no loss check), one 64-lane workgroup each.
  std::vector<int32_t> slots(2 * (size_t)(nelem + 1));
  std::vector<Instance> insts(nelem);
  rc = single_instance(lat, d_r, 6, 1, 12, 1, nullptr, 0, 0, 0, 1, nullptr, nullptr, nullptr,
                       nullptr, nullptr, 0, true, slots.data(), insts[0]);
  if (rc) {
    (void)hipFreeAsync(d_r, stream);
    return rc;
  }
  for (int e = 1; e < nelem; e++) {
    insts[e] = insts[0];
    insts[e].r = d_r + 72 * (size_t)e;
    insts[e].start_elem = e;
    insts[e].end_elem = e + 1;
  }
  const atx_launch_shape in{64, 1, 0, 0};
  atx_launch_shape shape;
  rc = resolve_shape(&in, shape, strict);
  hipError_t e = hipSuccess;
  if (!dev) e = hipMemcpyAsync(const_cast<double*>(d_o), orbits, no * sizeof(double),
                               hipMemcpyHostToDevice, stream);
  if (!rc && e == hipSuccess) e = atx_launch_element_probes(d_o, step, d_r, nelem, stream);
  if (!rc && e == hipSuccess)
    rc = launch_instances(insts, std::vector<const std::vector<int32_t>*>(nelem, &slots), shape,
                          stream, strict, lat->has_ext);
  if (!rc && e == hipSuccess) e = atx_launch_element_diffs(d_r, step, d_m, nelem, stream);
  if (!rc && e == hipSuccess && !dev)
    e = hipMemcpyAsync(m66, d_m, nm * sizeof(double), hipMemcpyDeviceToHost, stream);
  hipError_t fe = hipFreeAsync(d_r, stream);
  hipError_t se = dev ? hipSuccess : hipStreamSynchronize(stream);
  if (rc) return rc;
  if (e != hipSuccess) return fail(ATX_E_RUNTIME, std::string("element matrices: ") + hipGetErrorString(e));
  if (fe != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipFreeAsync: ") + hipGetErrorString(fe));
  if (se != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
  return ATX_OK;
}

int atx_elempass(atx_lattice* lat, int32_t index, double* r_in, uint32_t np, double energy,
                 double rest_energy, void* stream_, uint32_t flags) {
  g_err.clear();
  if (!lat) return fail(ATX_E_VALUE, "NULL lattice");
  std::lock_guard<std::mutex> g(lat->mu);
  const int nelem = (int)lat->elems.size();
  if (index < 0 || index >= nelem) return fail(ATX_E_VALUE, "element index out of range");
  if (np == 0) return ATX_OK;
  if (!r_in) return fail(ATX_E_VALUE, "r_in is NULL");
  int rc = ensure_device_lattice(lat, energy, rest_energy, true);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  const bool strict = (flags & ATX_STRICT) != 0;
  if (flags & ATX_DEVICE_POINTERS)
    return track_device(lat, r_in, 6, 1, np, 1, nullptr, 0, 0, index, index + 1, nullptr, nullptr,
                        nullptr, nullptr, nullptr, 0, true, stream, strict);
  double* d_r = nullptr;
  const size_t nr = 6 * (size_t)np;
  HIPCHK(hipMalloc(&d_r, nr * sizeof(double)));
  hipError_t e = hipMemcpy(d_r, r_in, nr * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    rc = track_device(lat, d_r, 6, 1, np, 1, nullptr, 0, 0, index, index + 1, nullptr, nullptr,
                      nullptr, nullptr, nullptr, 0, true, stream, strict);
    if (!rc) e = hipMemcpy(r_in, d_r, nr * sizeof(double), hipMemcpyDeviceToHost);
  }
  (void)hipFree(d_r);
  if (rc) return rc;
  if (e != hipSuccess) return fail(ATX_E_RUNTIME, std::string("hipMemcpy: ") + hipGetErrorString(e));
  return ATX_OK;
}

int atx_device_count(int* count) {
  if (!count) return fail(ATX_E_VALUE, "NULL argument");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return fail(ATX_E_NODEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  *count = n;
  return ATX_OK;
}

int atx_device_info(int device, char* name, int name_len, char* arch, int arch_len, int* cus,
                    char* platform, int platform_len) {
  hipDeviceProp_t p;
  hipError_t e = hipGetDeviceProperties(&p, device);
  if (e != hipSuccess) return fail(ATX_E_NODEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  int rv = 0;
  (void)hipRuntimeGetVersion(&rv);
  if (name && name_len > 0) std::snprintf(name, name_len, "%s", p.name);
  if (arch && arch_len > 0) std::snprintf(arch, arch_len, "%s", p.gcnArchName);
  if (cus) *cus = p.multiProcessorCount;
  if (platform && platform_len > 0)
    std::snprintf(platform, platform_len, "HIP %d.%d", rv / 10000000, (rv / 100000) % 100);
  return ATX_OK;
}

int atx_set_device(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(ATX_E_NODEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return ATX_OK;
}

int atx_get_device(int* device) {
  if (!device) return fail(ATX_E_VALUE, "NULL argument");
  HIPCHK(hipGetDevice(device));
  return ATX_OK;
}

// --- Device memory, pinned host memory, streams and events -----------------

int atx_device_alloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(ATX_E_VALUE, "NULL argument");
  *ptr = nullptr;
  if (bytes == 0) return ATX_OK;
  HIPCHK(hipMalloc(ptr, bytes));
  return ATX_OK;
}

int atx_device_free(void* ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return ATX_OK;
}

int atx_host_alloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(ATX_E_VALUE, "NULL argument");
  *ptr = nullptr;
  if (bytes == 0) return ATX_OK;
  HIPCHK(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return ATX_OK;
}

int atx_host_free(void* ptr) {
  if (ptr) HIPCHK(hipHostFree(ptr));
  return ATX_OK;
}

int atx_memcpy(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes == 0) return ATX_OK;
  if (!dst || !src) return fail(ATX_E_VALUE, "NULL argument");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, static_cast<hipStream_t>(stream)));
  return ATX_OK;
}

int atx_stream_create(void** stream) {
  if (!stream) return fail(ATX_E_VALUE, "NULL argument");
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return ATX_OK;
}

int atx_stream_destroy(void* stream) {
  if (stream) HIPCHK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return ATX_OK;
}

int atx_stream_synchronize(void* stream) {
  HIPCHK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return ATX_OK;
}

int atx_stream_wait_event(void* stream, void* event) {
  if (!event) return fail(ATX_E_VALUE, "NULL event");
  HIPCHK(hipStreamWaitEvent(static_cast<hipStream_t>(stream), static_cast<hipEvent_t>(event), 0));
  return ATX_OK;
}

int atx_event_create(void** event) {
  if (!event) return fail(ATX_E_VALUE, "NULL argument");
  hipEvent_t e = nullptr;
  HIPCHK(hipEventCreate(&e));
  *event = e;
  return ATX_OK;
}

int atx_event_destroy(void* event) {
  if (event) HIPCHK(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return ATX_OK;
}

int atx_event_record(void* event, void* stream) {
  if (!event) return fail(ATX_E_VALUE, "NULL event");
  HIPCHK(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return ATX_OK;
}

int atx_event_query(void* event, int* done) {
  if (!event || !done) return fail(ATX_E_VALUE, "NULL argument");
  hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
  if (e == hipErrorNotReady) {
    *done = 0;
    return ATX_OK;
  }
  HIPCHK(e);
  *done = 1;
  return ATX_OK;
}

int atx_event_synchronize(void* event) {
  if (!event) return fail(ATX_E_VALUE, "NULL event");
  HIPCHK(hipEventSynchronize(static_cast<hipEvent_t>(event)));
  return ATX_OK;
}

int atx_event_elapsed(float* ms, void* start, void* end) {
  if (!ms || !start || !end) return fail(ATX_E_VALUE, "NULL argument");
  HIPCHK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
  return ATX_OK;
}

}  // extern "C"
