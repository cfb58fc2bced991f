/* TEST INFRASTRUCTURE ONLY. Compiles the reference ExactDriftPass.c in
 * place (include path = the reference atintegrators directory) and exports
 * its static core under the pass name, for ref_driver.c. */
#include "ExactDriftPass.c"

void ExactDriftPass(double *r_in, double le, const double *T1, const double *T2,
                    const double *R1, const double *R2, double *RApertures,
                    double *EApertures, int num_particles)
{
    drift_pass(r_in, le, T1, T2, R1, R2, RApertures, EApertures, num_particles);
}
