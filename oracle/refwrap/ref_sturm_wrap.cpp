// Thin C wrapper around the reference's Sturm-sequence root finder
// (inc/maths/sturm_polynomial_solver.h, a std-only header), built from the
// reference source in place by oracle/Makefile into oracle/_ref/libref_sturm.so.
// TEST INFRASTRUCTURE ONLY: it pins the oracle's Sturm restatement, which
// the essential-matrix five-point solver uses for its degree-10 polynomial.
#include "maths/sturm_polynomial_solver.h"

extern "C" {

// coeffs: 11 ascending coefficients; roots: up to 10.  Returns the count.
int ref_bisect_sturm10(const double* coeffs, double* roots) {
  return poselib::sturm::bisect_sturm<10>(coeffs, roots, 1e-10);
}

}  // extern "C"
