// The reference's case files (see case_io.h).  Each reader follows the
// reference's READ statements record for record and each writer its WRITE
// statements, so that the files are interchangeable with the reference's.
#include "case_io.h"

#include <cmath>

namespace nemoh {
namespace casefiles {

using fio::edit_E;
using fio::edit_F;
using fio::edit_I;
using fio::FormatError;
using fio::Reader;
using fio::Writer;

namespace {

const double kPi = 4.0 * std::atan(1.0);  // PI=4.*ATAN(1.)

std::string join_path(const std::string& dir, const std::string& f) {
  if (f.empty() || f[0] == '/' || dir.empty() || dir == ".") return f;
  return dir.back() == '/' ? dir + f : dir + "/" + f;
}

DofCase read_dof(Reader& r, int body, int mode) {
  DofCase d;
  r.begin();
  d.icase = (int)r.integer();
  for (double& v : d.dir) v = r.real();
  for (double& v : d.axis) v = r.real();
  r.end();
  d.body = body;
  d.mode = mode;
  return d;
}

double abs_c(cplx z) { return std::hypot(z.real(), z.imag()); }
double arg_c(cplx z) { return std::atan2(z.imag(), z.real()); }

}  // namespace

// Nemoh.cal: environment (Environment.f90:34-42), bodies (Mesh.f90:120-143,
// Main.f90:164-184), load cases (Main.f90:111-131) and post-processing
// switches (Main.f90:132-147).  One pass reads the same records, in the
// same way, as the preProcessor's three passes.
NemohCal read_nemoh_cal(const std::string& path) {
  NemohCal c;
  Reader r(path);
  r.skip();
  r.begin();
  c.rho = r.real();
  r.end();
  r.begin();
  c.g = r.real();
  r.end();
  r.begin();
  c.depth = r.real();
  r.end();
  r.begin();
  c.xeff = r.real();
  c.yeff = r.real();
  r.end();
  r.skip();
  r.begin();
  const long long nbodies = r.integer();
  r.end();
  if (nbodies < 1) throw FormatError(path + ": the number of bodies must be at least 1");
  for (int b = 1; b <= nbodies; ++b) {
    BodyCal body;
    r.skip();
    r.begin();
    body.mesh_file = r.word();
    r.end();
    r.begin();
    body.npoints = (int)r.integer();
    body.npanels = (int)r.integer();
    r.end();
    r.begin();
    const long long nrad = r.integer();
    r.end();
    for (int i = 1; i <= nrad; ++i) body.rad.push_back(read_dof(r, b, i));
    r.begin();
    const long long nforce = r.integer();
    r.end();
    for (int i = 1; i <= nforce; ++i) body.force.push_back(read_dof(r, b, i));
    r.begin();
    const long long nadd = r.integer();
    r.end();
    for (int i = 0; i < nadd; ++i) r.skip();
    c.rad.insert(c.rad.end(), body.rad.begin(), body.rad.end());
    c.force.insert(c.force.end(), body.force.begin(), body.force.end());
    c.bodies.push_back(body);
  }
  r.skip();
  r.begin();
  c.nw = (int)r.integer();
  c.wmin = r.real();
  c.wmax = r.real();
  r.end();
  r.begin();
  c.nbeta = (int)r.integer();
  c.betamin = r.real();
  c.betamax = r.real();
  r.end();
  r.skip();
  r.skip();  // IRF line (read by the postProcessor)
  r.begin();
  c.switch_potential = (int)r.integer();
  r.end();
  r.begin();
  c.ntheta = (int)r.integer();
  c.thetamin = r.real();
  c.thetamax = r.real();
  r.end();
  r.begin();
  c.nx = (int)r.integer();
  c.ny = (int)r.integer();
  c.lx = r.real();
  c.ly = r.real();
  r.end();
  if (c.nw < 1 || c.nbeta < 0) throw FormatError(path + ": bad numbers of wave frequencies or directions");
  c.w.resize(c.nw);
  for (int j = 1; j <= c.nw; ++j)
    c.w[j - 1] = (c.nw > 1) ? c.wmin + (c.wmax - c.wmin) * (double)(j - 1) / (double)(c.nw - 1) : c.wmin;
  c.beta.resize(c.nbeta);
  for (int j = 1; j <= c.nbeta; ++j)
    c.beta[j - 1] = (c.nbeta > 1)
                        ? (c.betamin + (c.betamax - c.betamin) * (double)(j - 1) / (double)(c.nbeta - 1)) * kPi / 180.
                        : c.betamin * kPi / 180.;
  return c;
}

// Body meshes merged as the preProcessor reads them (Mesh.f90:144-195) and
// its panel geometry (Mesh.f90:196-226): two triangles P1P2P4 and P3P4P2.
PreMesh read_body_meshes(const NemohCal& cal, const std::string& nemoh_cal_path, const std::string& base_dir) {
  PreMesh m;
  m.nbodies = (int)cal.bodies.size();
  int npt = 0;
  for (size_t b = 0; b < cal.bodies.size(); ++b) {
    const BodyCal& body = cal.bodies[b];
    Reader r(join_path(base_dir, body.mesh_file));
    r.begin();
    (void)r.integer();
    const int isym = (int)r.integer();
    r.end();
    if (b > 0 && isym != m.isym)
      throw FormatError(" Error: there is an inconsistency in the mesh files regarding the xOz symmetries");
    m.isym = isym;
    for (int i = 0; i < body.npoints; ++i) {
      r.begin();
      (void)r.integer();
      const double x = r.real(), y = r.real(), z = r.real();
      r.end();
      const double tx = 0.0;  // Mesh.f90:150-157 adds tX (zero) to X and Y
      m.x.push_back(x + tx);
      m.y.push_back(y + tx);
      m.z.push_back(z);
    }
    r.skip();
    for (int i = 0; i < body.npanels; ++i) {
      r.begin();
      for (int k = 0; k < 4; ++k) m.p.push_back((int)r.integer() + npt);
      r.end();
      m.cpanel.push_back((int)b + 1);
    }
    npt += body.npoints;
  }
  (void)nemoh_cal_path;
  const int n = m.nfa();
  m.xm.resize(3 * (size_t)n);
  m.nrm.resize(3 * (size_t)n);
  m.a.resize(n);
  for (int i = 0; i < n; ++i) {
    double P[4][3];
    for (int l = 0; l < 4; ++l) {
      const int v = m.p[4 * (size_t)i + l] - 1;
      if (v < 0 || v >= m.np()) throw FormatError("a panel refers to a point outside the mesh");
      P[l][0] = m.x[v];
      P[l][1] = m.y[v];
      P[l][2] = m.z[v];
    }
    double u[3], v[3], w1[3], w2[3];
    for (int j = 0; j < 3; ++j) {
      u[j] = P[1][j] - P[0][j];
      v[j] = P[3][j] - P[1][j];
    }
    w1[0] = u[1] * v[2] - u[2] * v[1];
    w1[1] = u[2] * v[0] - u[0] * v[2];
    w1[2] = u[0] * v[1] - u[1] * v[0];
    const double a1 = 0.5 * std::sqrt(w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]);
    for (int j = 0; j < 3; ++j) {
      u[j] = P[3][j] - P[2][j];
      v[j] = P[1][j] - P[2][j];
    }
    w2[0] = u[1] * v[2] - u[2] * v[1];
    w2[1] = u[2] * v[0] - u[0] * v[2];
    w2[2] = u[0] * v[1] - u[1] * v[0];
    const double a2 = 0.5 * std::sqrt(w2[0] * w2[0] + w2[1] * w2[1] + w2[2] * w2[2]);
    const double a = a1 + a2;
    if (a < 1.0E-07)
      throw FormatError("Error: surface of panel " + edit_I(i + 1, 6) + " is too small (" + edit_E(a, 14, 7) + ")");
    m.a[i] = a;
    const double third = 1. / 3;
    for (int j = 0; j < 3; ++j)
      m.xm[3 * (size_t)i + j] =
          third * (P[0][j] + P[1][j] + P[3][j]) * a1 / a + third * (P[1][j] + P[2][j] + P[3][j]) * a2 / a;
    const double s[3] = {w1[0] + w2[0], w1[1] + w2[1], w1[2] + w2[2]};
    for (int j = 0; j < 3; ++j) m.nrm[3 * (size_t)i + j] = s[j] / std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  }
  return m;
}

// Mesh.f90:234-244
std::string l12_text(const PreMesh& m) {
  Writer w;
  w.line(edit_I(2, 3) + " " + edit_I(m.isym, 3));
  for (int i = 0; i < m.np(); ++i)
    w.line(edit_I(i + 1, 6) + " " + edit_E(m.x[i], 14, 7) + " " + edit_E(m.y[i], 14, 7) + " " + edit_E(m.z[i], 14, 7));
  w.line(edit_I(0, 6) + " " + edit_E(0., 14, 7) + " " + edit_E(0., 14, 7) + " " + edit_E(0., 14, 7));
  for (int i = 0; i < m.nfa(); ++i) {
    std::string l;
    for (int k = 0; k < 4; ++k) l += " " + edit_I(m.p[4 * (size_t)i + k], 6);
    w.line(l);
  }
  w.line(" " + edit_I(0, 6) + " " + edit_I(0, 6) + " " + edit_I(0, 6) + " " + edit_I(0, 6));
  return w.str();
}

// Mesh.f90:245-252
std::string l10_text(const PreMesh& m) {
  Writer w;
  w.begin().end();
  w.line(" " + edit_I(m.isym, 6) + " " + edit_I(m.np(), 6) + " " + edit_I(m.nfa(), 6) + " " + edit_I(m.nbodies, 6));
  for (int i = 0; i < m.nfa(); ++i) {
    std::string l = edit_I(m.cpanel[i], 5);
    for (int j = 0; j < 3; ++j) l += " " + edit_E(m.xm[3 * (size_t)i + j], 14, 7);
    for (int j = 0; j < 3; ++j) l += " " + edit_E(m.nrm[3 * (size_t)i + j], 14, 7);
    l += " " + edit_E(m.a[i], 14, 7);
    w.line(l);
  }
  w.begin().integer(0).end();
  return w.str();
}

// Mesh.f90:253-266
std::string mesh_tec_text(const PreMesh& m) {
  Writer w;
  w.line("VARIABLES=\"X\" \"Y\" \"Z\" \"NX\" \"NY\" \"NZ\" \"A\"");
  w.begin().text("ZONE N=").integer(m.np()).text(", E=").integer(m.nfa()).text(" , F=FEPOINT,ET=QUADRILATERAL").end();
  for (int i = 0; i < m.np(); ++i) {
    std::string l;
    for (double v : {m.x[i], m.y[i], m.z[i], 0.0, 0.0, 0.0, 0.0}) l += "  " + edit_E(v, 14, 7);
    w.line(l);
  }
  for (int i = 0; i < m.nfa(); ++i) {
    std::string l = edit_I(m.p[4 * (size_t)i], 6);
    for (int k = 1; k < 4; ++k) l += "  " + edit_I(m.p[4 * (size_t)i + k], 6);
    w.line(l);
  }
  w.begin().text("ZONE t=\"normales\", F=POINT, I=").integer(m.nfa()).end();
  for (int i = 0; i < m.nfa(); ++i) {
    std::string l;
    for (int j = 0; j < 3; ++j) l += "  " + edit_E(m.xm[3 * (size_t)i + j], 14, 7);
    for (int j = 0; j < 3; ++j) l += "  " + edit_E(m.nrm[3 * (size_t)i + j], 14, 7);
    l += "  " + edit_E(m.a[i], 14, 7);
    w.line(l);
  }
  return w.str();
}

// Main.f90:309-319
std::string kochin_dat_text(const NemohCal& c) {
  Writer w;
  w.begin().integer(c.ntheta).end();
  if (c.ntheta > 1) {
    for (int j = 1; j <= c.ntheta; ++j)
      w.begin().real((c.thetamin + (c.thetamax - c.thetamin) * (double)(j - 1) / (double)(c.ntheta - 1)) * kPi / 180.)
          .end();
  } else if (c.ntheta == 1) {
    w.begin().real(c.thetamin * kPi / 180.).end();
  }
  return w.str();
}

// Main.f90:293-306
std::string freesurface_dat_text(const NemohCal& c) {
  Writer w;
  w.begin().integer((long long)c.nx * c.ny).integer((long long)(c.nx - 1) * (c.ny - 1)).end();
  for (int i = 1; i <= c.nx; ++i)
    for (int j = 1; j <= c.ny; ++j)
      w.line(" " + edit_E(-0.5 * c.lx + c.lx * (double)(i - 1) / (double)(c.nx - 1), 14, 7) + " " +
             edit_E(-0.5 * c.ly + c.ly * (double)(j - 1) / (double)(c.ny - 1), 14, 7) + " " + edit_E(0., 14, 7));
  for (int i = 1; i <= c.nx - 1; ++i)
    for (int j = 1; j <= c.ny - 1; ++j)
      w.line(" " + edit_I(j + (i - 1) * c.ny, 7) + " " + edit_I(j + 1 + (i - 1) * c.ny, 7) + " " +
             edit_I(j + 1 + i * c.ny, 7) + " " + edit_I(j + i * c.ny, 7));
  return w.str();
}

// Main.f90:320-333 (the theta list has no NTheta = 1 case: 0/0 there)
std::string index_dat_text(const NemohCal& c) {
  Writer w;
  w.begin().integer(c.nw).integer(c.nbeta).integer((long long)c.rad.size()).integer((long long)c.force.size())
      .integer(c.ntheta).end();
  w.begin().text("--- Force ---").end();
  for (size_t k = 0; k < c.force.size(); ++k)
    w.begin().integer((long long)k + 1).integer(c.force[k].body).integer(c.force[k].mode).end();
  w.begin().text("--- Motion ---").end();
  for (size_t k = 0; k < c.rad.size(); ++k)
    w.begin().integer((long long)k + 1).integer(c.rad[k].body).integer(c.rad[k].mode).end();
  w.begin();
  for (double b : c.beta) w.real(b);
  w.end();
  w.begin();
  for (double x : c.w) w.real(x);
  w.end();
  w.begin();
  for (int k = 1; k <= c.ntheta; ++k)
    w.real((c.thetamin + (c.thetamax - c.thetamin) * (double)(k - 1) / (double)(c.ntheta - 1)) * kPi / 180.);
  w.end();
  return w.str();
}

// Main.f90:195-201
std::string integration_text(int nint, int ntot, const std::vector<double>& fnds) {
  Writer w;
  w.begin().integer(nint).end();
  for (int j = 0; j < nint; ++j) {
    w.begin();
    for (int c = 0; c < ntot; ++c) w.real(fnds[(size_t)j * ntot + c]);
    w.end();
  }
  return w.str();
}

// Main.f90:248-261
std::string normal_velocities_text(const NemohCal& c, int ntot, const std::vector<cplx>& nvel, int switch_fs,
                                   int switch_kochin) {
  const int per = c.nbeta + (int)c.rad.size(), nprob = per * c.nw;
  Writer w;
  w.begin().integer(nprob).end();
  w.begin();
  for (int i = 0; i < c.nw; ++i)
    for (int j = 0; j < per; ++j) w.real(c.w[i]);
  w.end();
  for (int i = 0; i < c.nw; ++i) {
    w.begin();
    for (double b : c.beta) w.real(b);
    for (size_t j = 0; j < c.rad.size(); ++j) w.real(-1.);
    w.end();
  }
  for (int sw : {c.switch_potential, switch_fs, switch_kochin}) {
    w.begin();
    for (int q = 0; q < nprob; ++q) w.integer(sw);
    w.end();
  }
  for (int cc = 0; cc < ntot; ++cc) {
    w.begin();
    for (int q = 0; q < nprob; ++q) {
      const cplx v = nvel[(size_t)q * ntot + cc];
      w.real(v.real()).real(v.imag());
    }
    w.end();
  }
  return w.str();
}

// Main.f90:276-280
std::string fkforce_dat_text(const NemohCal& c, const std::vector<cplx>& fk) {
  const int nint = (int)c.force.size();
  Writer w;
  for (int k = 0; k < nint; ++k) {
    w.begin();
    for (int i = 0; i < c.nw; ++i) {
      for (int j = 0; j < c.nbeta; ++j) {
        const cplx f = fk[((size_t)i * c.nbeta + j) * nint + k];
        w.real(abs_c(f)).real(arg_c(f));
      }
      for (size_t j = 1; j <= c.rad.size(); ++j) w.real(0. * (double)j).real(0. * (double)j);
    }
    w.end();
  }
  return w.str();
}

// Main.f90:264-275
std::string fkforce_tec_text(const NemohCal& c, const std::vector<cplx>& fk) {
  const int nint = (int)c.force.size();
  Writer w;
  w.line("VARIABLES=\"w (rad/s)\"");
  for (int k = 0; k < nint; ++k)
    w.line("\"abs(F" + edit_I(c.force[k].body, 4) + edit_I(k + 1, 4) + ")\" \"angle(F" + edit_I(c.force[k].body, 4) +
           edit_I(k + 1, 4) + ")\"");
  for (int j = 0; j < c.nbeta; ++j) {
    w.line("Zone t=\"FKforce - beta = " + edit_F(c.beta[j] * 180. / kPi, 7, 3) + "\",I=" + edit_I(c.nw, 6) +
           ",F=POINT");
    for (int i = 0; i < c.nw; ++i) {
      std::string l = " " + edit_E(c.w[i], 14, 7);
      for (int k = 0; k < nint; ++k) {
        const cplx f = fk[((size_t)i * c.nbeta + j) * nint + k];
        l += " " + edit_E(abs_c(f), 14, 7) + " " + edit_E(arg_c(f), 14, 7);
      }
      w.line(l);
    }
  }
  return w.str();
}

Environment read_environment(const std::string& path) {
  Environment e;
  Reader r(path);
  r.skip();
  r.begin();
  e.rho = r.real();
  r.end();
  r.begin();
  e.g = r.real();
  r.end();
  r.begin();
  e.depth = r.real();
  r.end();
  r.begin();
  e.xeff = r.real();
  e.yeff = r.real();
  r.end();
  return e;
}

SolverMesh read_solver_mesh(const std::string& dir) {
  SolverMesh m;
  int npoints, npanels;
  {
    Reader r(join_path(dir, "mesh/L10.dat"));
    r.skip();
    r.begin();
    (void)r.integer();
    npoints = (int)r.integer();
    npanels = (int)r.integer();
    m.nbodies = (int)r.integer();
    r.end();
  }
  {
    Reader r(join_path(dir, "mesh/L12.dat"));
    r.begin();
    const long long fmt = r.integer();
    m.isym = (int)r.integer();
    r.end();
    if (fmt != 2) throw FormatError(" The mesh file format is not correct. ");
    for (int i = 0; i < npoints; ++i) {
      r.begin();
      (void)r.integer();
      m.x.push_back(r.real());
      m.y.push_back(r.real());
      m.z.push_back(r.real());
      r.end();
    }
    r.skip();
    for (int i = 0; i < npanels; ++i) {
      r.begin();
      for (int k = 0; k < 4; ++k) m.p.push_back((int)r.integer());
      r.end();
    }
  }
  {
    Reader r(join_path(dir, "mesh/L10.dat"));
    r.skip();
    r.begin();
    const long long isym = r.integer();
    r.end();
    if (isym != m.isym) throw FormatError(" The mesh file format is not correct. ");
    for (int i = 0; i < npanels; ++i) {
      r.begin();
      m.cpanel.push_back((int)r.integer());
      for (int k = 0; k < 7; ++k) (void)r.real();
      r.end();
    }
  }
  return m;
}

BodyConditions read_body_conditions(const std::string& path, int ntot) {
  BodyConditions b;
  Reader r(path);
  r.begin();
  b.nprob = (int)r.integer();
  r.end();
  const int n = b.nprob;
  b.omega.resize(n);
  b.type.resize(n);
  b.sw_potential.resize(n);
  b.sw_fs.resize(n);
  b.sw_kochin.resize(n);
  r.begin();
  for (double& v : b.omega) v = r.real();
  r.end();
  r.begin();
  for (double& v : b.type) v = r.real();
  r.end();
  for (std::vector<int>* s : {&b.sw_potential, &b.sw_fs, &b.sw_kochin}) {
    r.begin();
    for (int& v : *s) v = (int)r.integer();
    r.end();
  }
  b.nvel.assign((size_t)n * ntot, cplx(0, 0));
  for (int c = 0; c < ntot; ++c) {
    r.begin();
    for (int q = 0; q < n; ++q) {
      const double re = r.real(), im = r.real();
      b.nvel[(size_t)q * ntot + c] = cplx(re, im);
    }
    r.end();
  }
  return b;
}

std::vector<double> read_integration(const std::string& path, int ntot, int* nint) {
  Reader r(path);
  r.begin();
  *nint = (int)r.integer();
  r.end();
  std::vector<double> nds((size_t)*nint * ntot);
  for (int j = 0; j < *nint; ++j) {
    r.begin();
    for (int c = 0; c < ntot; ++c) nds[(size_t)j * ntot + c] = r.real();
    r.end();
  }
  return nds;
}

std::vector<double> read_kochin_dat(const std::string& path) {
  Reader r(path);
  r.begin();
  const long long n = r.integer();
  r.end();
  std::vector<double> t;
  for (long long j = 0; j < n; ++j) {
    r.begin();
    t.push_back(r.real());
    r.end();
  }
  return t;
}

FreeSurfaceMesh read_freesurface_dat(const std::string& path) {
  FreeSurfaceMesh fs;
  Reader r(path);
  r.begin();
  fs.npoints = (int)r.integer();
  fs.npanels = (int)r.integer();
  r.end();
  if (fs.npoints > 0) {
    for (int j = 0; j < fs.npoints; ++j) {
      r.begin();
      fs.x.push_back(r.real());
      fs.y.push_back(r.real());
      r.end();
    }
    for (int j = 0; j < fs.npanels; ++j) {
      r.begin();
      for (int k = 0; k < 4; ++k) fs.p.push_back((int)r.integer());
      r.end();
    }
  }
  return fs;
}

// NEMOH.f90:150-165
std::string forces_text(const BodyConditions& bc, int nint, const std::vector<cplx>& force) {
  Writer w;
  for (int c = 0; c < nint; ++c) {
    w.begin();
    for (int j = 0; j < bc.nprob; ++j) {
      const cplx f = force[(size_t)j * nint + c];
      if (bc.type[j] != -1.)
        w.real(abs_c(f)).real(arg_c(f));
      else
        w.real(f.imag() / bc.omega[j]).real(-f.real());
    }
    w.end();
  }
  return w.str();
}

std::string kochin_text(const std::vector<double>& theta, const cplx* h) {
  Writer w;
  for (size_t i = 0; i < theta.size(); ++i) w.begin().real(theta[i]).real(abs_c(h[i])).real(arg_c(h[i])).end();
  return w.str();
}

std::string freesurface_text(const FreeSurfaceMesh& fs, const cplx* eta) {
  Writer w;
  w.begin().text("VARIABLES=\"X\" \"Y\" \"abs(eta) (m)\" \"angle(phi) (rad)\" \"PRE1\" \"PRE2\"").end();
  w.line("ZONE N=" + edit_I(fs.npoints, 7) + " ,E = " + edit_I(fs.npanels, 7) + " ,F=FEPOINT,ET=QUADRILATERAL");
  for (int i = 0; i < fs.npoints; ++i)
    w.line(" " + edit_E(fs.x[i], 14, 7) + " " + edit_E(fs.y[i], 14, 7) + " " + edit_E(abs_c(eta[i]), 14, 7) + " " +
           edit_E(arg_c(eta[i]), 14, 7) + " " + edit_E(eta[i].real(), 14, 7) + " " + edit_E(eta[i].imag(), 14, 7));
  for (int i = 0; i < fs.npanels; ++i) {
    w.begin();
    for (int k = 0; k < 4; ++k) w.integer(fs.p[4 * (size_t)i + k]);
    w.end();
  }
  return w.str();
}

std::string pressure_text(const SolverMesh& m, const cplx* pressure) {
  const int nfa = m.nfa(), nsym = (m.isym == 1) ? 1 : 0, nb = 1 << nsym;
  Writer w;
  w.begin().text("VARIABLES=\"X\" \"Y\" \"Z\" \"abs(p) (Pa)\" \"angle(p) (rad)\"").end();
  w.line("ZONE N=" + edit_I(4 * nfa * nb, 7) + ",E = " + edit_I(nfa * nb, 7) + ",F=FEPOINT,ET=QUADRILATERAL");
  for (int s = 0; s < nb; ++s)
    for (int i = 0; i < nfa; ++i) {
      const cplx p = pressure[i + (size_t)s * nfa];
      for (int k = 0; k < 4; ++k) {
        const int v = m.p[4 * (size_t)i + k] - 1;
        w.begin().real(m.x[v]).real(s ? -m.y[v] : m.y[v]).real(m.z[v]).real(abs_c(p)).real(arg_c(p)).end();
      }
    }
  for (int s = 0; s < nb; ++s)
    for (int i = 1; i <= nfa; ++i) {
      const int o = (i - 1) * 4 + s * 4 * nfa;
      w.begin().integer(1 + o).integer(2 + o).integer(3 + o).integer(4 + o).end();
    }
  return w.str();
}

std::string problem_tag(int number) { return edit_I(number, 5); }

}  // namespace casefiles
}  // namespace nemoh
