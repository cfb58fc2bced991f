/* TEST INFRASTRUCTURE ONLY. Compiles the reference ExactMultipoleRadPass.c
 * in place and exports its static core under the pass name, for
 * ref_driver.c. */
#include "ExactMultipoleRadPass.c"

void ExactMultipoleRadPass(double *r, double le, double *A, double *B, int max_order,
                           int num_int_steps, int FringeQuadEntrance, int FringeQuadExit,
                           double *T1, double *T2, double *R1, double *R2,
                           double *RApertures, double *EApertures, double *KickAngle,
                           double scaling, double gamma, int num_particles, double *bdiff)
{
    multipole_pass(r, le, A, B, max_order, num_int_steps, FringeQuadEntrance, FringeQuadExit,
                   T1, T2, R1, R2, RApertures, EApertures, KickAngle, scaling, gamma,
                   num_particles, bdiff);
}
