// Per-mesh and per-frequency scalar work of the solver core, run on the host:
// panel geometry (PREPARE_MESH.f90:30-180), the theta rule (CINT_*), the
// finite-depth wave number (X0) and the exponential (Prony) fit of the
// finite-depth kernel (LISC and helpers, COMPUTE_GREEN_FD.f90:1007-1680).
//
// These are O(N) or O(1) per frequency and stay on the host next to the
// kernel launches.  The Prony fit is numerically ill-conditioned at small
// kh, so it keeps the reference's exact sequence of operations (including
// how the compiler runtime divides complex numbers) to reproduce the
// reference's fitted amplitudes bit for bit.
#include "host_math.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace nemoh {
namespace host {

using cd = std::complex<double>;

int prepare_mesh(int np, int nfa, int nsym, const double* x, const double* y, const double* z, const int* m1,
                 const int* m2, const int* m3, const int* m4, PanelGeometry* out) {
  double gl = 0.0, bl = 0.0;
  for (int i = 0; i < np; ++i)
    for (int j = i + 1; j < np; ++j) {
      gl = std::max(std::fabs(x[j] - x[i]), gl);
      if (nsym == 0) bl = std::max(std::fabs(y[j] - y[i]), bl);
      else bl = std::max(std::max(std::fabs(y[j]), std::fabs(y[i])), bl);
    }
  if (nsym == 1) bl = 2.0 * bl;
  out->zer = -1.E-5 * std::max(gl, bl);
  for (int i = 0; i < nfa; ++i) {
    const int k = m1[i], l = m2[i], m = m3[i], n = m4[i];
    const double t1x = x[m] - x[k], t1y = y[m] - y[k], t1z = z[m] - z[k];
    const double t2x = x[n] - x[l], t2y = y[n] - y[l], t2z = z[n] - z[l];
    const double nx = t1y * t2z - t2y * t1z, ny = t1z * t2x - t2z * t1x, nz = t1x * t2y - t2x * t1y;
    const double nq = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (nq <= 1.E-25) return i + 1;
    const double xn = nx / nq, yn = ny / nq, zn = nz / nq;
    out->xn[i] = xn;
    out->yn[i] = yn;
    out->zn[i] = zn;
    // vertex average in the reference's summation order
    const double xa = (x[k] + x[m] + x[l] + x[n]) * 0.25;
    const double ya = (y[k] + y[m] + y[n] + y[l]) * 0.25;
    const double za = (z[k] + z[m] + z[n] + z[l]) * 0.25;
    const double t1 = std::sqrt(t1x * t1x + t1y * t1y + t1z * t1z);
    const double ux = t1x / t1, uy = t1y / t1, uz = t1z / t1;  // in-plane axis 1
    const double vx = yn * uz - zn * uy, vy = zn * ux - xn * uz, vz = xn * uy - yn * ux;  // axis 2
    auto xi = [&](int p) { return ux * (x[p] - xa) + uy * (y[p] - ya) + uz * (z[p] - za); };
    auto eta = [&](int p) { return vx * (x[p] - xa) + vy * (y[p] - ya) + vz * (z[p] - za); };
    const double xi1 = xi(k), xi2 = xi(l), xi3 = xi(m), xi4 = xi(n);
    const double e1 = eta(k), e2 = eta(l), e3 = eta(m), e4 = eta(n);
    // collocation point: centroid of the quadrilateral in its own plane
    const double xio = (xi4 * (e1 - e2) + xi2 * (e4 - e1)) / (3.0 * (e2 - e4));
    const double etao = -e1 / 3.0;
    out->xg[i] = xa + ux * xio + vx * etao;
    out->yg[i] = ya + uy * xio + vy * etao;
    out->zg[i] = za + uz * xio + vz * etao;
    const double a1 = xi1 - xio, a2 = xi2 - xio, a3 = xi3 - xio, a4 = xi4 - xio;
    const double b1 = e1 - etao, b2 = e2 - etao, b3 = e3 - etao, b4 = e4 - etao;
    out->aire[i] = (a3 - a1) * (b4 - b2) * 0.5;
    out->tdis[i] = std::max(std::max(std::sqrt(a1 * a1 + b1 * b1), std::sqrt(a2 * a2 + b2 * b2)),
                            std::max(std::sqrt(a3 * a3 + b3 * b3), std::sqrt(a4 * a4 + b4 * b4)));
  }
  return 0;
}

// Gauss half-rules on [-1/2, 1/2] as tabulated by the reference (7 digits).
namespace {
const double kQ8[] = {.4801449, .3983332, .2627662, .09171732};
const double kW8[] = {.05061427, .1111905, .1568533, .1813419};
const double kQ12[] = {.4907803, .4520586, .3849513, .2936589, .1839157, .06261670};
const double kW12[] = {.2358766E-1, .5346966E-1, .8003916E-1, .1015837, .1167462, .1245735};
const double kQ16[] = {.4947004, .4722875, .4328156, .3777022, .3089381, .2290084, .1408017, .04750625};
const double kW16[] = {.01357622, .03112676, .04757925, .06231448, .07479799, .08457826, .09130170, .09472530};
const double kQ24[] = {.4975936, .4873642, .469137,  .4432077, .4100009, .3700621,
                       .3240468, .2727107, .2168967, .1575213, .09555943, .032028446};
const double kW24[] = {.6170615E-2, .1426569E-1, .2213872E-1, .2964929E-1, .366732E-1,  .4309508E-1,
                       .4880932E-1, .5372213E-1, .5775283E-1, .6083523E-1, .6291873E-1, .6396909E-1};
const double kQ32[] = {.4986319, .4928057, .4823811, .4674530, .4481605, .4246838, .3972418, .3660910,
                       .3315221, .2938578, .2534499, .2106756, .1659343, .1196436, .07223598, .02415383};
const double kW32[] = {.350930E-2,  .8137197E-2, .1269603E-1, .1713693E-1, .2141794E-1, .2549903E-1,
                       .2934204E-1, .3291111E-1, .3617289E-1, .3909694E-1, .4165596E-1, .4382604E-1,
                       .4558693E-1, .4692219E-1, .4781936E-1, .4827004E-1};
}  // namespace

int theta_rule(double akk, double* qq, double* cq) {
  const double pi = 4.0 * std::atan(1.0);
  const double *q, *w;
  int n;
  if (akk <= .4) { n = 8; q = kQ8; w = kW8; }
  else if (akk <= 2.5) { n = 12; q = kQ12; w = kW12; }
  else if (akk <= 4.) { n = 16; q = kQ16; w = kW16; }
  else if (akk <= 8.) { n = 24; q = kQ24; w = kW24; }
  else if (akk <= 25.) { n = 32; q = kQ32; w = kW32; }
  else {  // Simpson on [-pi/2, pi/2]
    n = (akk > 40.) ? 101 : 51;
    for (int j = 1; j <= n; ++j) {
      qq[j - 1] = -pi / 2. + (j - 1.) / (n - 1.) * pi;
      if (j == 1 || j == n) cq[j - 1] = pi / (3. * (n - 1.));
      else cq[j - 1] = ((j % 2) ? 2. : 4.) / (3. * (n - 1.)) * pi;
    }
    return n;
  }
  for (int i = 0; i < n / 2; ++i) {
    qq[i] = q[i] * pi;
    qq[n - 1 - i] = -q[i] * pi;
    cq[i] = cq[n - 1 - i] = w[i] * pi;
  }
  return n;
}

// Wavenumber (Common/Environment.f90:45-84): deep water (w^2/g) when k h is
// outside (0, 20], else a bisection of x tanh x = k h to a relative width of
// 1e-6 (its iteration counter is reset before the second loop, as there).
double incident_wavenumber(double w, double depth, double g) {
  double k = w * w / g;
  const double kh = k * depth;
  if (kh <= 20.0 && kh > 0.0) {
    double xg = 0.0, xd = kh;
    int nite = 0;
    while (nite < 10000 && (kh - xg * std::tanh(xg)) * (kh - xd * std::tanh(xd)) > 0.0) {
      xg = xd;
      xd = 2.0 * xd;
      ++nite;
    }
    nite = 0;
    double xc = 0.5 * (xd + xg);
    while (nite < 10000 && std::fabs(xd - xg) / std::fabs(xc) >= 1.0e-06) {
      xc = 0.5 * (xd + xg);
      if ((kh - xg * std::tanh(xg)) * (kh - xc * std::tanh(xc)) > 0.0)
        xg = xc;
      else
        xd = xc;
      ++nite;
    }
    if (nite >= 10000) return std::nan("");
    k = xc / depth;
  }
  return k;
}

double x0(double ak) {
  const double eps = 5.E-6;
  double xi = 0.0, xs = 0.0;
  const double pas = std::max(ak, std::sqrt(ak));
  int itour = 0;
  do {  // bracket
    xs += pas;
    ++itour;
  } while (itour <= 1000 && (ak - xs * std::tanh(xs)) * (ak - xi * std::tanh(xi)) > 0);
  for (int iter = 1;; ++iter) {  // bisection, with the reference's 1e-12 clamp
    const double xm = (xi + xs) * 0.5;
    if (iter > 1000 || itour > 1000) return NAN;
    if (!(std::fabs((xs - xi) / xm) > eps)) return xm;
    double v1 = ak - xm * std::tanh(xm);
    const double v2 = ak - xi * std::tanh(xi);
    if (std::fabs(v1 * v2) <= 1.E-12) v1 = 0.0;
    if (v1 * v2 < 0) xs = xm;
    else xi = xm;
  }
}

// ---------------------------------------------------------------------------
// Prony fit.  The finite-depth kernel FF(x) (COMPUTE_GREEN_FD.f90:1612) is
// sampled on [-0.1, 20]; a linear-prediction least-squares problem gives the
// characteristic polynomial, whose roots (Bairstow) give the exponents; a
// second least-squares problem gives the amplitudes; the number of terms
// grows 4, 6, ..., 30 until the fit is within 1e-2.
namespace {

double kernel_ff(double x, double ak, double am) {
  const double coef = (am + ak) * (am + ak) / (am * am - ak * ak + ak);
  const double tol = std::max(0.1, 0.1 * am);
  auto f = [&](double t) { return (t + ak) * std::exp(t) / (t * std::sinh(t) - ak * std::cosh(t)) - coef / (t - am) - 2.; };
  if (std::fabs(x - am) > tol) return f(x);
  // quadratic through (am - tol, am, am + tol) across the removable pole
  const double a = am - tol, b = am, c = am + tol;
  const double fa = f(a), fc = f(c);
  const double fb = coef / (am + ak) * (am + ak + 1) - (coef / (am + ak)) * (coef / (am + ak)) * am - 2;
  return (x - b) * (x - c) * fa / ((a - b) * (a - c)) + (x - c) * (x - a) * fb / ((b - c) * (b - a)) +
         (x - a) * (x - b) * fc / ((c - a) * (c - b));
}

// Householder least squares; column-major matrix with leading dimension 120
// and (ncc + 1) columns, the last being the right-hand side.
struct LsMatrix {
  std::vector<double> a = std::vector<double>(120 * 32, 0.0);
  double& operator()(int i, int j) { return a[(i - 1) + 120 * (j - 1)]; }
};

bool householder_ls(LsMatrix& A, int nl, int ncc, double eps) {
  if (ncc > nl) return false;
  const int ntc = ncc + 1;
  for (int k = 1; k <= ncc; ++k) {
    double e = 0;
    for (int i = k; i <= nl; ++i) e = e + A(i, k) * A(i, k);
    const double e0 = std::sqrt(e);
    if (e0 < eps) return false;
    const double ar = (A(k, k) == 0) ? -e0 : -std::copysign(e0, A(k, k));
    const double eta = ar * (ar - A(k, k));
    for (int j = k + 1; j <= ntc; ++j) {
      double ba = (A(k, k) - ar) * A(k, j);
      for (int i = k + 1; i <= nl; ++i) ba = ba + A(i, k) * A(i, j);
      A(k, j) = A(k, j) + ba / ar;
      for (int i = k + 1; i <= nl; ++i) A(i, j) = A(i, j) - A(i, k) * ba / eta;
    }
    A(k, k) = ar;
    for (int i = k + 1; i <= nl; ++i) A(i, k) = 0;
  }
  // no unknowns (LISC kept no decaying exponent): nothing to back-substitute.
  // HOUSRS (COMPUTE_GREEN_FD.f90:1454) indexes A(0, .) here, out of bounds.
  if (ncc < 1) return true;
  A(ncc, ntc) = A(ncc, ntc) / A(ncc, ncc);
  for (int i1 = ncc - 1; i1 >= 1; --i1) {
    for (int i = i1 + 1; i <= ncc; ++i) A(i1, ntc) = A(i1, ntc) - A(i1, i) * A(i, ntc);
    A(i1, ntc) = A(i1, ntc) / A(i1, i1);
  }
  return true;
}

// One quadratic factor x^2 + q1 x + q2 of the polynomial c[1..ic] by
// Bairstow iteration (COMPUTE_GREEN_FD.f90:1320-1451).  The control flow is
// the reference's: convergence tests, restarts and error codes.
int bairstow_factor(double* c, int ic, double* q, int lim) {
  double h = 0., hh = 0., a = 0, a1 = 0, aa = 0, b = 0, bb = 0, b1 = 0, c1 = 0, ca = 0, cb = 0, cc = 0, cd = 0;
  double q1 = 0, q2 = 0, qq1 = 0, qq2 = 0, qqq1 = 0, qqq2 = 0, dq1 = 0, dq2 = 0;
  const double eps = 1.E-14, eps1 = 1.E-6;
  int j = ic, l = 0, ll = 0, i;
  while (j > 1 && c[j] == 0) --j;
  if (j <= 1) return -1;
  a = c[j];
  if (a != 1.) {
    for (i = 1; i <= j; ++i) {
      c[i] = c[i] / a;
      if (std::fabs(c[i]) >= 1.E27) return -1;
    }
  }
  if (j < 3) return -2;
  if (j == 3) {
    q[1] = c[1];
    q[2] = c[2];
    q[3] = 0.;
    q[4] = 0.;
    return 0;
  }
  enum { kRestart, kIterate } state = kRestart;
  int ier = 0;
  for (;;) {
    if (state == kRestart) {
      l = 0;
      ll = 0;
      q1 = q[1];
      q2 = q[2];
      qq1 = 0.;
      qq2 = 0.;
      aa = c[1];
      bb = c[2];
      cb = std::fabs(aa);
      ca = std::fabs(bb);
      if (cb < ca) {
        cc = cb + cb;
        cb = cb / ca;
        ca = 1.;
      } else if (cb == ca) {
        cc = ca + ca;
        ca = 1.;
        cb = 1.;
      } else {
        cc = ca + ca;
        ca = ca / cb;
        cb = 1.;
      }
      cd = cc * .1;
      state = kIterate;
    }
    // synthetic division by the current factor
    a = 0.;
    b = 0.;
    a1 = 0.;
    b1 = 0.;
    i = j;
    qqq1 = q1;
    qqq2 = q2;
    dq1 = hh;
    dq2 = h;
    bool overflow = false;
    for (;;) {
      h = -q1 * b - q2 * a + c[i];
      if (std::fabs(h) >= 1.E27) { overflow = true; break; }
      b = a;
      a = h;
      --i;
      if (i < 1) break;
      if (i == 1) h = 0.;
      h = -q1 * b1 - q2 * a1 + h;
      if (std::fabs(h) >= 1.E27) { overflow = true; break; }
      c1 = b1;
      b1 = a1;
      a1 = h;
    }
    if (overflow) { ier = -3; break; }
    h = ca * std::fabs(a) + cb * std::fabs(b);
    if (ll > 0) {
      q[1] = q1;
      q[2] = q2;
      q[3] = a;
      q[4] = b;
      return 0;
    }
    ++l;
    if (std::fabs(a) <= eps * std::fabs(c[1]) && std::fabs(b) <= eps * std::fabs(c[2])) {
      q[1] = q1;
      q[2] = q2;
      q[3] = a;
      q[4] = b;
      return 0;
    }
    if (h <= cc) {
      aa = a;
      bb = b;
      cc = h;
      qq1 = q1;
      qq2 = q2;
    }
    if (l > lim) {
      if (h <= cd) { ier = 1; break; }
      if (q[1] == 0 && q[2] == 0) { ier = -3; break; }
      q[1] = 0.;
      q[2] = 0.;
      state = kRestart;
      continue;
    }
    hh = std::max(std::max(std::fabs(a1), std::fabs(b1)), std::fabs(c1));
    if (hh <= 0) { ier = -3; break; }
    a1 = a1 / hh;
    b1 = b1 / hh;
    c1 = c1 / hh;
    h = a1 * c1 - b1 * b1;
    if (h == 0) { ier = -3; break; }
    a = a / hh;
    b = b / hh;
    hh = (b * a1 - a * b1) / h;
    h = (a * c1 - b * b1) / h;
    q1 = q1 + hh;
    q2 = q2 + h;
    if (std::fabs(hh) <= eps * std::fabs(q1) && std::fabs(h) <= eps * std::fabs(q2)) {
      ll = 1;
      continue;
    }
    if (l <= 1) continue;
    if (std::fabs(hh) > eps1 * std::fabs(q1) || std::fabs(h) > eps1 * std::fabs(q2)) continue;
    if (std::fabs(qqq1 * hh) >= std::fabs(q1 * dq1)) break;  // ier = 0, best so far
    if (std::fabs(qqq2 * h) < std::fabs(q2 * dq2)) continue;
    break;
  }
  q[1] = qq1;
  q[2] = qq2;
  q[3] = aa;
  q[4] = bb;
  return ier;
}

// All roots of the polynomial c[1..ic] (ascending powers) by repeated
// quadratic deflation (COMPUTE_GREEN_FD.f90:1182-1317).  Returns the error
// code (2 is fatal in the reference).
int polynomial_roots(double* c, int ic, double* rr, double* rc) {
  double pol[33] = {0};
  const double eps = 1.E-6;
  const int lim = 100;
  int ir = ic, ist = 0, j, ier = 0;
  while (ir > 1 && c[ir] == 0) --ir;
  if (ir <= 1) return 2;
  j = ir;
  int l = 0;
  double a = c[ir], b, h, q1, q2, q[5];
  for (int i = 1; i <= ir; ++i) {
    if (l <= 0) {
      if (c[i] == 0) {  // zero root
        rr[i] = 0.;
        rc[i] = 0.;
        pol[j] = 0.;
        --j;
        continue;
      }
      l = 1;
      ist = i;
      j = 0;
    }
    ++j;
    c[i] = c[i] / a;
    pol[j] = c[i];
    if (std::fabs(pol[j]) >= 1.E27) return 2;
  }
  q1 = 0.;
  q2 = 0.;
  for (;;) {
    if (j < 2) {
      --ir;
      break;
    }
    if (j == 2) {  // last linear factor
      a = pol[1];
      rr[ist] = -a;
      rc[ist] = 0.;
      --ir;
      q2 = 0.;
      for (int i = 2; i <= ir; ++i) {
        q1 = q2;
        q2 = pol[i + 1];
        pol[i] = a * q2 + q1;
      }
      pol[ir + 1] = a + q2;
      break;
    }
    int found = 0;
    for (l = 1; l <= 10 && !found; ++l) {
      for (int n = 1;; ++n) {
        q[1] = q1;
        q[2] = q2;
        const int code = bairstow_factor(pol, j, q, lim);
        if (code >= 0) {
          found = 1;
          if (code > 0) ier = 1;
          break;
        }
        if ((q1 != 0 || q2 != 0) && n <= 3) {  // retry with sign flips of the guess
          if (n == 2) q2 = -q2;
          else q1 = -q1;
          continue;
        }
        q1 = 1. + q1;
        q2 = 1. - q2;
        break;
      }
    }
    if (!found) return 3;
    q1 = q[1];
    q2 = q[2];
    // deflate
    b = 0.;
    a = 0.;
    int i = j;
    do {
      h = -q1 * b - q2 * a + pol[i];
      pol[i] = b;
      b = a;
      a = h;
      --i;
    } while (i > 2);
    pol[2] = b;
    pol[1] = a;
    l = ir - 1;
    for (i = j; i <= l; ++i) pol[i - 1] = pol[i - 1] + pol[i] * q2 + pol[i + 1] * q1;
    pol[l] = pol[l] + pol[l + 1] * q2 + q1;
    pol[ir] = pol[ir] + q2;
    // roots of x^2 + q1 x + q2
    h = -.5 * q2;
    a = h * h - q1;
    b = std::sqrt(std::fabs(a));
    if (a <= 0) {
      rr[ist] = h;
      rc[ist] = b;
      rr[ist + 1] = h;
      rc[ist + 1] = -b;
    } else {
      b = h + std::copysign(b, h);
      rr[ist] = q1 / b;
      rc[ist] = 0.;
      rr[ist + 1] = b;
      rc[ist + 1] = 0.;
    }
    ist += 2;
    j -= 2;
  }
  a = 0.;
  for (int i = 1; i <= ir; ++i) {
    q1 = c[i];
    q2 = pol[i + 1];
    pol[i] = q2;
    if (q1 != 0) q2 = (q1 - q2) / q1;
    a = std::max(a, std::fabs(q2));
  }
  rr[ir + 1] = a;
  rc[ir + 1] = 0.;
  if (ier <= 0 && a > eps) ier = -1;
  return ier;
}

// complex / real the way the compiler runtime divides two complex numbers
// (operands scaled by a power of two, then (a c + b d) / (c^2 + d^2)); the
// reference's CLOG(COM)/H goes through this path.
cd div_scaled(cd num, double h) {
  const double a = num.real(), b = num.imag();
  double c = h, d = 0.0;
  int e = 0;
  const double lb = std::logb(std::fmax(std::fabs(c), std::fabs(d)));
  if (std::isfinite(lb)) {
    e = (int)lb;
    c = std::scalbn(c, -e);
    d = std::scalbn(d, -e);
  }
  const double den = c * c + d * d;
  return cd(std::scalbn((a * c + b * d) / den, -e), std::scalbn((b * c - a * d) / den, -e));
}

}  // namespace

int prony_fit(double akh, double amh, int* nexp_out, double* ambda, double* ar) {
  const double lo = -0.1, hi = 20.;
  double xt[122] = {0}, yt[122] = {0};
  double amb[33] = {0}, amp[33] = {0};
  for (int k = 0; k < 31; ++k) amb[k + 1] = ambda[k];
  int nm = 0;
  for (int nk = 4;; nk += 2) {
    nm = nk;
    const int nj = 4 * nm, npp = nj + 1;
    const double h = (hi - lo) / nj;
    for (int i = 1; i <= npp; ++i) {
      xt[i] = lo + (i - 1) * h;
      yt[i] = kernel_ff(xt[i], akh, amh);
    }
    // linear prediction: sum_j s_j y(m - j + i) = -y(m + i)
    LsMatrix S;
    const int k = npp - nm;
    for (int i = 1; i <= k; ++i) {
      for (int j = 1; j <= nm; ++j) S(i, j) = yt[nm - j + i];
      S(i, nm + 1) = -yt[nm + i];
    }
    if (!householder_ls(S, k, nm, 1.E-20)) return 1;
    double sc[33] = {0}, vr[33] = {0}, vc[33] = {0};
    for (int i = 1; i <= nm; ++i) sc[nm - i + 1] = S(i, nm + 1);
    sc[nm + 1] = 1.;
    if (polynomial_roots(sc, nm + 1, vr, vc) == 2) return 2;
    const double hs = (xt[npp] - xt[1]) / nj;
    for (int i = 1; i <= nm; ++i) {
      const cd com = div_scaled(std::log(cd(vr[i], vc[i])), hs);
      vr[i] = com.real();
      vc[i] = com.imag();
    }
    // one exponent per real root or complex-conjugate pair, keeping the
    // decaying ones above -20
    int cnt = 0;
    for (int i = 1; i <= nm;) {
      amb[++cnt] = vr[i];
      i += (vc[i] != 0 && std::fabs(vr[i] - vr[i + 1]) <= 1.E-5) ? 2 : 1;
    }
    int kept = 0;
    for (int i = 1; i <= cnt; ++i)
      if (amb[i] < 0. && amb[i] + 20. > 0) amb[++kept] = amb[i];
    nm = kept;
    // amplitudes: normal equations of sum_j ar_j exp(l_j x) ~ y, by Householder
    LsMatrix A;
    for (int i = 1; i <= nm; ++i)
      for (int j = 1; j <= nm; ++j) {
        double s = 0;
        for (int p = 1; p <= npp; ++p) {
          const double t = (amb[i] + amb[j]) * xt[p];
          if (t + 30 >= 0) s = s + std::exp(t);
        }
        A(i, j) = s;
      }
    for (int i = 1; i <= nm; ++i) {
      double s = 0;
      for (int p = 1; p <= npp; ++p) {
        const double t = amb[i] * xt[p];
        if (t + 30 >= 0) s = s + std::exp(t) * yt[p];
      }
      A(i, nm + 1) = s;
    }
    if (!householder_ls(A, nm, nm, 1.E-20)) return 3;
    for (int i = 1; i <= nm; ++i) amp[i] = A(i, nm + 1);
    // fit check on a grid twice as fine over [0, 20]
    const int nmo = 2 * npp - 1;
    bool ok = true;
    for (int i = 1; i <= nmo; ++i) {
      const double xx = (i - 1) * hi / (nmo - 1);
      double tt = 0.;
      for (int j = 1; j <= nm; ++j) {
        const double rt = amb[j] * xx;
        if (rt > -20.) tt = tt + amp[j] * std::exp(rt);
      }
      if (std::fabs(kernel_ff(xx, akh, amh) - tt) > 1.E-02) ok = false;
    }
    if (ok || nk + 2 > 30) break;
  }
  for (int j = 1; j <= nm; ++j)
    if (amb[j] > 0.) return 4;
  for (int k = 0; k < 31; ++k) {
    ambda[k] = amb[k + 1];
    ar[k] = amp[k + 1];
  }
  *nexp_out = nm;
  return 0;
}

}  // namespace host
}  // namespace nemoh
