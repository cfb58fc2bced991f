// atx_device.h -- compiled-lattice format shared by the host compiler
// (atx_host.cpp) and the tracking kernels (atx_kernel.hip).
//
// A lattice is compiled into an array of fixed 64-byte element headers plus
// one pool of doubles holding every element-constant value. Everything that
// depends only on element parameters, the ring energy or the ring length is
// evaluated once on the host in FP64 with the C library (the same values the
// reference integrators compute inside their particle loops: KickAngle
// folding, Forest-Ruth slice lengths, irho, edge tan(), elegant fringe
// matrices, rad_const = RAD_CONST*gamma^3, RF nv/2*pi*f/(h/f - T0)). The
// kernel reads headers and pool through uniform (scalar) loads: every lane of
// a workgroup walks the same element sequence.
#pragma once

#include <stdint.h>

namespace atx {

// XCDs of the device: the hardware deals the workgroups of a launch to them
// in turn (workgroup i to XCD i mod kXcds), and each has its own L2.
constexpr int kXcds = 8;

// Dispatch kinds of the kernel switch (one per in-scope pass method).
enum Kind : int32_t {
  K_IDENTITY = 0,
  K_DRIFT = 1,
  K_STR = 2,      // StrMPoleSymplectic4Pass
  K_BND = 3,      // BndMPoleSymplectic4Pass
  K_STRRAD = 4,   // StrMPoleSymplectic4RadPass
  K_BNDRAD = 5,   // BndMPoleSymplectic4RadPass
  K_THIN = 6,     // ThinMPolePass
  K_RF = 7,       // RFCavityPass
  K_CORR = 8,     // CorrectorPass
  K_APERTURE = 9, // AperturePass
  K_CAVITY = 10,      // CavityPass
  K_EXDRIFT = 11,     // ExactDriftPass
  K_EXMPOLE = 12,     // ExactMultipolePass
  K_EXMPOLERAD = 13   // ExactMultipoleRadPass
};

// Header flags.
enum : uint32_t {
  F_EDGEB_IN = 1u << 20,   // FringeInt1 > 0 and FullGap > 0 (diff_bend_fringe.c:36-44)
  F_EDGEB_OUT = 1u << 21,  // FringeInt2 > 0 and FullGap > 0
  F_T1 = 1u << 0,
  F_R1 = 1u << 1,
  F_R2 = 1u << 2,
  F_T2 = 1u << 3,
  F_RAP = 1u << 4,
  F_EAP = 1u << 5,
  F_SCALE = 1u << 6,      // FieldScaling != 1
  F_THICK = 1u << 7,      // RF / corrector with non-zero length
  F_RFKICK = 1u << 8,     // RF with nv != 0
};

// Edge-focusing modes (vertical term fy) chosen on the host from
// FringeBendEntrance/Exit and whether the fringe correction is zero.
enum EdgeMode : int32_t {
  E_CONST = 0,   // fy = irho*tan(edge)                         (element constant)
  E_DIVP = 1,    // fy = (irho*tan(edge)) / (1+dp)               non-rad method 2, corr = 0
  E_MULPN = 2,   // fy = (irho*tan(edge)) * pn                   rad method 2, psi = 0
  E_TAN1 = 3,    // fy = irho*tan(edge - corr/(1+dp))            non-rad method 1/legacy
  E_TAN2 = 4,    // fy = irho*tan(edge - corr/(1+dp))/(1+dp)     non-rad method 2
  E_TAN3 = 5,    // fy = irho*tan(edge - corr +- px/(1+dp))      non-rad method 3
  E_RTAN1 = 6,   // fy = irho*tan(edge - psi*pn)                 rad method 1/legacy
  E_RTAN2 = 7,   // fy = irho*tan(edge - psi*pn)*pn              rad method 2
  E_RTAN3 = 8    // fy = irho*tan(edge - psi +- px*pn)           rad method 3
};

// Quadrupole-fringe modes.
enum QuadFringe : int32_t { Q_NONE = 0, Q_LEE = 1, Q_ELEGANT = 2 };

// Scalar-block slots (pool[h.sc + S_*]).
enum Slot : int32_t {
  S_L1 = 0, S_L2 = 1, S_K1 = 2, S_K2 = 3,   // Forest-Ruth drift/kick lengths
  S_IRHO = 4,
  S_SCALE = 5, S_INVSCALE = 6,
  S_B2 = 7,                                 // PolynomB[1] for the quad fringes
  S_FX_IN = 8, S_FY_IN = 9, S_FX_OUT = 10, S_FY_OUT = 11,
  S_EDGE_IN = 12, S_CORR_IN = 13, S_EDGE_OUT = 14, S_CORR_OUT = 15,
  S_RADC = 16,
  S_BAX = 17, S_BAY = 18,
  S_RK1 = 19, S_RK2 = 20,                   // S_RADC * S_K1, S_RADC * S_K2 (fast build)
  S_ELEN = 21,                              // Length (exact multipoles)
  S_DIFC = 22,                              // DIF_CONST * gamma^5 (diffusion matrices)
  S_HL1 = 23, S_HL2 = 24,                   // S_L1 / 2, S_L2 / 2 (fast radiating drifts)
  S_NSCALAR = 26,
  // RF cavity
  S_RF_HALF = 0, S_RF_NV = 1, S_RF_TWOPIF = 2, S_RF_LAG = 3, S_RF_PHILAG = 4, S_RF_HFT0 = 5,
  // corrector
  S_C_XK = 0, S_C_YK = 1, S_C_LEN = 2, S_C_Q0 = 3,  // S_C_Q0 = xk*xk/3 + yk*yk/3
  // drift
  S_D_LEN = 0
};

struct DevElem {
  int32_t kind;
  uint32_t flags;
  int32_t order;      // MaxOrder
  int32_t nslice;     // NumIntSteps
  int32_t sc;         // pool offset of the scalar block
  int32_t poly;       // pool offset of A[0..order], then B[0..order]
  int32_t t1, r1;     // pool offsets (valid when the flag is set)
  int32_t r2, t2;
  int32_t rap, eap;   // RApertures (4) / Limits (4); EApertures (2)
  int32_t edge;       // EdgeMode at entrance | (EdgeMode at exit << 8); DriftPass: low
                      // word of the Length bits
  int32_t ops;        // pool offset of the slice op table (integrators other than 4), -1: none
  int32_t fringe;     // QuadFringe at entrance | (QuadFringe at exit << 8); DriftPass: high
                      // word of the Length bits
  int32_t fq;         // pool offset of the elegant fringe matrices: in-P, in-M, out-M, out-P (8 each)
};

// Slice op table of a non-default integrator: pool[ops] = number of ops,
// then (code, length) pairs with code 0 = drift of that length, 1 = kick.
static_assert(sizeof(DevElem) == 64, "DevElem must stay 64 bytes (dword fields: scalar loads)");

// One tracked lattice instance inside a launch. A single-lattice call has one
// instance; batched error realisations have one per lattice.
struct Instance {
  const DevElem* elems;
  const double* pool;
  const int32_t* slots;     // 2*(nelem+1): refpt copy rank per element, first turn then later turns
  double* r;                // particle state, updated in place
  double* rout;             // refpt copies (6, N, R, T) F-order
  uint8_t* lost_flag;
  uint32_t* lost_turn;
  uint32_t* lost_elem;
  double* lost_coord;       // (6, N) F-order
  int64_t pstride;          // distance between particles in r (6: (6,N) F-order; 1: SoA)
  int64_t cstride;          // distance between coordinates in r (1: F-order; N: SoA)
  int32_t nelem;
  int32_t npart;
  int32_t nturns;
  int32_t turn0;
  int32_t start_elem;
  int32_t end_elem;
  int32_t n_first;          // copies made on the first turn
  int32_t n_std;            // copies made on each later turn
  int32_t block_begin;      // first workgroup of this instance
  int32_t losses;           // 1: record the loss map (at.c checkiflost); 0: setlost only;
                            // 2: as 1, continuing the map of an earlier call
  int32_t elempass;         // 1: single element pass, no loss check (at.c at_elempass)
  int32_t pad;
  // Repacked launches: lanes take particles idx[0..*nlive) instead of
  // 0..npart (both written by repack_kernel); NULL for plain launches.
  const int32_t* idx;
  const int32_t* nlive;
};

// Per-element data of the adjoint kernel: gradient slot (-1: none) and
// PolynomB index of the differentiated coefficient; pool offsets of the
// inverse misalignment matrices.
struct AdjElem {
  int32_t slot;
  int32_t coeff;
  int32_t r1inv;
  int32_t r2inv;
};

// Diffusion matrices (one lane per element): B of each radiating element
// from its entrance orbit.
struct DiffArgs {
  const DevElem* elems;
  const double* pool;
  const double* orbits;  // (6, nelem): orbit at the entrance of each element
  double* bdiff;         // nelem x 36, column-major 6x6 per element
  int32_t nelem;
  int32_t pad;
};

// One block of turns of the adjoint sweep (gridDim.y indexes the items).
struct AdjItem {
  const double* states;  // (nturns+1) blocks of (6, N): turn-start states, then the end state
  double* lam;           // (6, N) adjoint of the end state in, of the start state out
  double* grad;          // accumulated dL/dPolynomB[coeff] per slot; NULL: not accumulated
  int32_t npart;
  int32_t nturns;
  int32_t turn0;         // turn counter of the first turn of the block
  int32_t pad;
};

struct AdjArgs {
  const DevElem* elems;
  const double* pool;
  const AdjElem* adj;
  const AdjItem* items;
  int32_t nelem;
  int32_t nitems;
};

}  // namespace atx
