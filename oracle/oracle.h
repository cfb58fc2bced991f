/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU oracle for the MI355X tracking engine. Nothing in the product path
 * (at_amd/) may include, link or call this; only tests/, the smoke check
 * and the CPU-baseline leg of bench.py use it, as the checker.
 *
 * One flat element record describes every in-scope pass method. It carries
 * the pass-method parameters under the reference attribute names
 * (pyat/at/lattice/elements/element_object.py:46-63 and the trackFunction
 * getters of each atintegrators/<Pass>.c). Pointers are NULL when the
 * optional attribute is absent, exactly as atGetOptionalDoubleArray returns
 * NULL (atintegrators/atelem.c:336-352).
 *
 * Two implementations share this interface:
 *   oracle.c      -- a C restatement of the reference algorithm (this repo);
 *   ref_driver.c  -- a driver restating pyat/at.c:596-654 that calls the
 *                    reference integrator cores compiled in place from
 *                    /root/reference/atintegrators (built into oracle/_ref/).
 */
#ifndef ATX_ORACLE_H
#define ATX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pass-method type IDs, same numbering as atgpu/Element.h:14-28. */
enum {
    OR_IDENTITYPASS = 0,
    OR_DRIFTPASS = 1,
    OR_BNDMPOLESYMPLECTIC4PASS = 2,
    OR_BNDMPOLESYMPLECTIC4RADPASS = 3,
    OR_STRMPOLESYMPLECTIC4PASS = 4,
    OR_STRMPOLESYMPLECTIC4RADPASS = 5,
    OR_CAVITYPASS = 6,
    OR_RFCAVITYPASS = 7,
    OR_EXACTDRIFTPASS = 8,
    OR_EXACTMULTIPOLEPASS = 9,
    OR_EXACTMULTIPOLERADPASS = 10,
    OR_THINMPOLEPASS = 11,
    OR_CORRECTORPASS = 12,
    OR_APERTUREPASS = 13
};

typedef struct oracle_elem {
    int32_t pass;
    int32_t MaxOrder;
    int32_t NumIntSteps;
    int32_t FringeBendEntrance;
    int32_t FringeBendExit;
    int32_t FringeQuadEntrance;
    int32_t FringeQuadExit;
    int32_t pad0;
    double Length;
    double BendingAngle;
    double EntranceAngle;
    double ExitAngle;
    double FringeInt1;
    double FringeInt2;
    double FullGap;
    double FieldScaling;
    double Voltage;
    double Frequency;
    double TimeLag;
    double PhaseLag;
    double bax;            /* ThinMPolePass BendingAngle[0] */
    double bay;            /* ThinMPolePass BendingAngle[1] */
    double Energy;         /* element Energy (CavityPass), 0 when absent */
    const double *PolynomA;
    const double *PolynomB;
    const double *fringeIntM0;
    const double *fringeIntP0;
    const double *R1;
    const double *R2;
    const double *T1;
    const double *T2;
    const double *RApertures;
    const double *EApertures;
    const double *KickAngle;
    const double *Limits;
} oracle_elem;

/*
 * Multi-turn tracking with the semantics of pyat/at.c:at_atpass
 * (:340-681): r_in is (6,N) F-order and is updated in place; r_out receives
 * the (6,N,R,T) F-order refpt copies in the sequential order at.c writes
 * them; with losses != 0 the loss map arrays are filled (survivors 0).
 * Returns 0 on success, otherwise a non-zero code; oracle_last_error()
 * gives the message.
 */
int oracle_atpass(const oracle_elem *elems, int32_t nelem, double *r_in,
                  int32_t num_particles, int32_t nturns,
                  const uint32_t *refpts, int32_t num_refpts, int32_t turn,
                  double energy, double rest_energy, int32_t losses,
                  int32_t start_elem, int32_t end_elem, double *r_out,
                  uint8_t *lost_flag, uint32_t *lost_turn, uint32_t *lost_elem,
                  double *lost_coord, int32_t nthreads);

/* Single-element pass with the semantics of pyat/at.c:at_elempass
 * (:683-753): nturn = 0, T0 = 0, no loss check. */
int oracle_elempass(const oracle_elem *elem, double *r_in, int32_t num_particles,
                    double energy, double rest_energy);

const char *oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
