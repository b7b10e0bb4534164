// Host-side scalar routines of the solver core (see host_math.cpp).
#pragma once

namespace nemoh {
namespace host {

struct PanelGeometry {
  double *xg, *yg, *zg, *xn, *yn, *zn, *aire, *tdis;  // [nfa] outputs
  double zer;
};

// Collocation points, normals, areas, TDIS and ZER; 0-based connectivity.
// Returns 0 or the 1-based index of a degenerate panel.
int prepare_mesh(int np, int nfa, int nsym, const double* x, const double* y, const double* z, const int* m1,
                 const int* m2, const int* m3, const int* m4, PanelGeometry* out);
// Theta rule for AKK = k * GM; returns the rule length.
int theta_rule(double akk, double* qq, double* cq);
// Root of x tanh x = ak by the reference's bisection.
double x0(double ak);
// The preProcessor's wave number k(w, depth, g) (NaN if its bisection fails).
double incident_wavenumber(double w, double depth, double g);
// Exponential fit of the finite-depth kernel; returns 0 or an error code.
int prony_fit(double akh, double amh, int* nexp, double* ambda, double* ar);

}  // namespace host
}  // namespace nemoh
