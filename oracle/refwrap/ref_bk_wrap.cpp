// Thin C wrapper around the reference's own graph-cut code (energy.h,
// graph.h, graph.cpp, maxflow.cpp, block.h: Boykov-Kolmogorov max-flow), built
// from the reference sources in place by oracle/Makefile into
// oracle/_ref/libref_bk.so.  TEST INFRASTRUCTURE ONLY: it pins the oracle's
// exact-min-cut labelling against the reference implementation.
//
// The call sequence mirrors GCRANSAC::labeling (inc/GCRANSAC.h:913-1021):
// add_node per point, add_term1 per point, add_term2 per same-cell pair p<q
// in the caller's order, minimize, what_segment.
#include "energy.h"
#include "graph.cpp"
#include "maxflow.cpp"

extern "C" {

// unary_a/unary_b: E(x=0), E(x=1) per node; edges (ep, eq) with terms
// (A, B, C, D).  Writes 1 for SINK.  Returns the minimised energy.
double ref_bk_minimize(int n, const double* unary_a, const double* unary_b, long ne, const int* ep, const int* eq,
                       const double* ta, const double* tb, const double* tc, const double* td, unsigned char* sink) {
  typedef Energy<double, double, double> E;
  E* g = new E(n, (int)(ne > 0 ? ne : 1), nullptr);
  for (int i = 0; i < n; ++i) g->add_node();
  for (int i = 0; i < n; ++i) g->add_term1(i, unary_a[i], unary_b[i]);
  for (long e = 0; e < ne; ++e) g->add_term2(ep[e], eq[e], ta[e], tb[e], tc[e], td[e]);
  const double energy = g->minimize();
  for (int i = 0; i < n; ++i) sink[i] = g->what_segment(i) == Graph<double, double, double>::SINK ? 1 : 0;
  delete g;
  return energy;
}

// The usage example of energy.h:31-69 (E(x,y,z) = x - 2y + 3(1-z) - 4xy + 5|y-z|).
double ref_bk_energy_example(int* x, int* y, int* z) {
  typedef Energy<double, double, double> E;
  E* e = new E(3, 2, nullptr);
  E::Var vx = e->add_variable(), vy = e->add_variable(), vz = e->add_variable();
  e->add_term1(vx, 0, 1);
  e->add_term1(vy, 0, -2);
  e->add_term1(vz, 3, 0);
  e->add_term2(vx, vy, 0, 0, 0, -4);
  e->add_term2(vy, vz, 0, 5, 5, 0);
  const double emin = e->minimize();
  *x = e->get_var(vx);
  *y = e->get_var(vy);
  *z = e->get_var(vz);
  delete e;
  return emin;
}
}
