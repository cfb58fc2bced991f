// Host-side self-check of the C ABI under AddressSanitizer (no GPU needed):
// builds lattices with every pass method and attribute kind, compiles them
// (atx_lattice_check runs the whole host compiler), walks the error paths,
// and destroys the handles. Exit status 0 when every expectation holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/atx.h"

static int failures = 0;
#define EXPECT(cond)                                                        \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::fprintf(stderr, "FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond, \
                   atx_last_error());                                       \
      failures++;                                                           \
    }                                                                       \
  } while (0)

static void arr(atx_lattice* l, const char* name, std::vector<double> v) {
  EXPECT(atx_element_set_array(l, name, v.data(), (int64_t)v.size()) == ATX_OK);
}

static void magnet(atx_lattice* l, const char* pass, double len, int order) {
  EXPECT(atx_lattice_add_element(l, pass, "M") == ATX_OK);
  atx_element_set_double(l, "Length", len);
  // thick magnets read PolynomB[1] (quadrupole fringes) whatever MaxOrder is
  const size_t n = (size_t)(order + 1 > 2 ? order + 1 : 2);
  std::vector<double> a(n, 0.0), b(n, 0.0);
  for (size_t k = 0; k < n; k++) b[k] = 0.1 * (double)(k + 1);
  arr(l, "PolynomA", a);
  arr(l, "PolynomB", b);
  atx_element_set_int(l, "MaxOrder", order);
  atx_element_set_int(l, "NumIntSteps", 5);
}

static void misalign(atx_lattice* l) {
  std::vector<double> t{1e-5, 0, -1e-5, 0, 0, 0}, r(36, 0.0);
  for (int i = 0; i < 6; i++) r[i * 7] = 1.0;
  arr(l, "T1", t);
  arr(l, "T2", t);
  arr(l, "R1", r);
  arr(l, "R2", r);
  arr(l, "RApertures", {-0.01, 0.01, -0.005, 0.005});
  arr(l, "EApertures", {0.01, 0.005});
}

int main() {
  for (int round = 0; round < 3; round++) {
    atx_lattice* l = nullptr;
    EXPECT(atx_lattice_create(&l) == ATX_OK);
    EXPECT(atx_lattice_add_element(l, "IdentityPass", "I") == ATX_OK);
    misalign(l);
    EXPECT(atx_lattice_add_element(l, "DriftPass", "D") == ATX_OK);
    atx_element_set_double(l, "Length", 1.5);
    misalign(l);
    for (const char* p : {"StrMPoleSymplectic4Pass", "StrMPoleSymplectic4RadPass",
                          "ExactMultipolePass", "ExactMultipoleRadPass"}) {
      for (int order = 0; order <= 6; order++) magnet(l, p, 0.4, order);
      atx_element_set_int(l, "FringeQuadEntrance", 1);
      atx_element_set_double(l, "FieldScaling", 1.01);
      arr(l, "KickAngle", {1e-5, -1e-5});
      misalign(l);
    }
    for (const char* p : {"BndMPoleSymplectic4Pass", "BndMPoleSymplectic4RadPass"}) {
      for (int m = 0; m <= 3; m++) {
        magnet(l, p, 1.0, 1);
        atx_element_set_double(l, "BendingAngle", 0.05);
        atx_element_set_double(l, "EntranceAngle", 0.02);
        atx_element_set_double(l, "ExitAngle", 0.01);
        atx_element_set_int(l, "FringeBendEntrance", m);
        atx_element_set_int(l, "FringeBendExit", m);
        atx_element_set_double(l, "FullGap", 0.03);
        atx_element_set_double(l, "FringeInt1", 0.5);
        atx_element_set_double(l, "FringeInt2", 0.4);
        atx_element_set_int(l, "FringeQuadEntrance", 2);
        arr(l, "fringeIntM0", {0.01, 0.02, 0.003, 0.0004, 0.005});
        arr(l, "fringeIntP0", {0.011, -0.02, 0.003, -0.0004, 0.005});
      }
    }
    magnet(l, "ThinMPolePass", 0.0, 3);
    arr(l, "BendingAngle", {1e-3, 5e-4});
    EXPECT(atx_lattice_add_element(l, "RFCavityPass", "RF") == ATX_OK);
    atx_element_set_double(l, "Length", 0.5);
    atx_element_set_double(l, "Voltage", 6e6);
    atx_element_set_double(l, "Frequency", 3.52e8);
    atx_element_set_int(l, "HarmNumber", 992);
    atx_element_set_double(l, "Energy", 6e9);
    EXPECT(atx_lattice_add_element(l, "CavityPass", "C") == ATX_OK);
    atx_element_set_double(l, "Length", 0.0);
    atx_element_set_double(l, "Voltage", 2e6);
    atx_element_set_double(l, "Frequency", 3.52e8);
    atx_element_set_double(l, "Energy", 6e9);
    EXPECT(atx_lattice_add_element(l, "CorrectorPass", "K") == ATX_OK);
    atx_element_set_double(l, "Length", 0.2);
    arr(l, "KickAngle", {1e-4, -2e-4});
    EXPECT(atx_lattice_add_element(l, "AperturePass", "A") == ATX_OK);
    arr(l, "Limits", {-0.01, 0.01, -0.005, 0.005});
    EXPECT(atx_lattice_add_element(l, "ExactDriftPass", "X") == ATX_OK);
    atx_element_set_double(l, "Length", 0.7);
    EXPECT(atx_lattice_num_elements(l) == 44);
    for (int integ = 1; integ <= 6; integ++) {
      EXPECT(atx_lattice_set_integrator(l, integ) == ATX_OK);
      EXPECT(atx_lattice_check(l, 6e9, 0.51099895e6) == ATX_OK);
    }
    EXPECT(atx_lattice_set_integrator(l, 7) != ATX_OK);
    EXPECT(atx_lattice_set_integrator(l, 4) == ATX_OK);
    // change one attribute after compiling, then compile again
    double kb[2] = {0.0, 0.7};
    EXPECT(atx_element_set_array_at(l, 3, "PolynomB", kb, 2) == ATX_OK);
    EXPECT(atx_element_set_array_at(l, 100000, "PolynomB", kb, 2) != ATX_OK);
    EXPECT(atx_lattice_check(l, 6e9, 0.51099895e6) == ATX_OK);
    int64_t turn = -1;
    EXPECT(atx_lattice_turn_counter(l, &turn) == ATX_OK && turn == 0);
    EXPECT(atx_lattice_destroy(l) == ATX_OK);

    // error paths
    atx_lattice* e = nullptr;
    EXPECT(atx_lattice_create(&e) == ATX_OK);
    EXPECT(atx_lattice_add_element(e, "NoSuchPass", "Z") != ATX_OK);
    EXPECT(std::strstr(atx_last_error(), "NoSuchPass") != nullptr);
    EXPECT(atx_lattice_add_element(e, "DriftPass", "D") == ATX_OK);
    EXPECT(atx_lattice_check(e, 3e9, 0.0) != ATX_OK);  // Length missing
    atx_element_set_double(e, "Length", 1.0);
    EXPECT(atx_lattice_check(e, 3e9, 0.0) == ATX_OK);
    EXPECT(atx_lattice_add_element(e, "StrMPoleSymplectic4Pass", "Q") == ATX_OK);
    atx_element_set_double(e, "Length", 0.4);
    EXPECT(atx_lattice_check(e, 3e9, 0.0) != ATX_OK);  // PolynomB missing
    arr(e, "PolynomA", {0.0, 0.0});
    arr(e, "PolynomB", {0.0, 1.0});
    atx_element_set_int(e, "MaxOrder", 5);  // beyond the polynomials
    atx_element_set_int(e, "NumIntSteps", 10);
    EXPECT(atx_lattice_check(e, 3e9, 0.0) == ATX_OK || atx_last_error()[0] != 0);
    arr(e, "R1", {1.0, 2.0});  // wrong size
    EXPECT(atx_lattice_check(e, 3e9, 0.0) != ATX_OK);
    EXPECT(atx_element_set_array(e, "T1", nullptr, 6) != ATX_OK);
    EXPECT(atx_element_set_double(nullptr, "Length", 1.0) != ATX_OK);
    // adjoint item lists are validated before any device work
    atx_adjoint_item it{};
    it.nturns = 1;
    it.num_particles = 10;
    EXPECT(atx_adjoint_batch(e, &it, 1, 3e9, 0.0, nullptr, nullptr, nullptr, 0, nullptr) != ATX_OK);
    EXPECT(atx_adjoint_batch(e, nullptr, 2, 3e9, 0.0, nullptr, nullptr, nullptr, 0, nullptr) != ATX_OK);
    EXPECT(atx_adjoint_batch(e, &it, 70000, 3e9, 0.0, nullptr, nullptr, nullptr, 0, nullptr) != ATX_OK);
    EXPECT(atx_adjoint_batch(nullptr, &it, 1, 3e9, 0.0, nullptr, nullptr, nullptr, 0, nullptr) != ATX_OK);
    EXPECT(atx_lattice_destroy(e) == ATX_OK);
    EXPECT(atx_lattice_destroy(nullptr) == ATX_OK);
  }
  if (failures) std::fprintf(stderr, "%d failures\n", failures);
  else std::printf("host self-check passed\n");
  return failures ? 1 : 0;
}
