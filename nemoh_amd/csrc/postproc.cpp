// The reference postProcessor (postProcessor/Main.f90:21-101): hydrodynamic
// coefficients from the Solver's files (Common/Results.f90:132-258) and the
// radiation impulse response functions (postProcessor/IRF.f90:66-151).
// Plain host arithmetic in the reference's order of operations; the inputs
// are the text files, so these outputs do not depend on where the solve ran.
//
// Not reproduced: Compute_RAOs (postProcessor/Compute_RAOs.f90:21-42) only
// zeroes its array, and Plot_WaveElevation runs under a switch that
// Initialize_Plot_WaveElevation never sets (Plot_WaveElevation.f90:116-133).
#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "case_io.h"

namespace nemoh {
namespace casefiles {

using fio::format_write;
using fio::FormatError;
using fio::FmtItem;
using fio::Reader;

namespace {

struct Results {  // TResults, Common/Results.f90:22-37
  int nw = 0, nbeta = 0, nrad = 0, nint = 0, ntheta = 0;
  std::vector<double> w, beta, theta;
  std::vector<int> force_body, force_mode, rad_body, rad_mode;
  // (i, j, k) = (frequency, heading or motion, force): [(i * n2 + j) * nint + k]
  std::vector<cplx> diffraction, fk;
  std::vector<double> added_mass, damping;
  size_t d(int i, int j, int k) const { return ((size_t)i * nbeta + j) * nint + k; }
  size_t r(int i, int j, int k) const { return ((size_t)i * nrad + j) * nint + k; }
};

std::string in_dir(const std::string& dir, const std::string& rel) {
  if (dir.empty() || dir == ".") return rel;
  return dir.back() == '/' ? dir + rel : dir + "/" + rel;
}

// amplitude * EXP(i phase), as line(c)*CEXP(CMPLX(0.,1.)*line(c+1))
cplx polar_(double a, double ph) { return cplx(a * std::cos(ph), a * std::sin(ph)); }
double abs_c(cplx z) { return std::hypot(z.real(), z.imag()); }
double arg_c(cplx z) { return std::atan2(z.imag(), z.real()); }

// ReadTResults, Common/Results.f90:132-190
Results read_results(const std::string& dir) {
  Results R;
  {
    Reader r(in_dir(dir, "results/index.dat"));
    r.begin();
    R.nw = (int)r.integer();
    R.nbeta = (int)r.integer();
    R.nrad = (int)r.integer();
    R.nint = (int)r.integer();
    R.ntheta = (int)r.integer();
    r.end();
    r.skip();
    for (int k = 0; k < R.nint; ++k) {
      r.begin();
      (void)r.integer();
      R.force_body.push_back((int)r.integer());
      R.force_mode.push_back((int)r.integer());
      r.end();
    }
    r.skip();
    for (int k = 0; k < R.nrad; ++k) {
      r.begin();
      (void)r.integer();
      R.rad_body.push_back((int)r.integer());
      R.rad_mode.push_back((int)r.integer());
      r.end();
    }
    for (auto* v : {&R.beta, &R.w, &R.theta}) {
      const int n = (v == &R.beta) ? R.nbeta : (v == &R.w) ? R.nw : R.ntheta;
      r.begin();
      for (int k = 0; k < n; ++k) v->push_back(r.real());
      r.end();
    }
  }
  R.diffraction.assign((size_t)R.nw * R.nbeta * R.nint, cplx(0, 0));
  R.fk.assign(R.diffraction.size(), cplx(0, 0));
  R.added_mass.assign((size_t)R.nw * R.nrad * R.nint, 0.0);
  R.damping.assign(R.added_mass.size(), 0.0);
  {
    Reader r(in_dir(dir, "results/Forces.dat"));
    const int per = 2 * R.nw * (R.nbeta + R.nrad);
    std::vector<double> line(per);
    for (int k = 0; k < R.nint; ++k) {
      r.begin();
      for (double& v : line) v = r.real();
      r.end();
      int c = 0;
      for (int i = 0; i < R.nw; ++i) {
        for (int j = 0; j < R.nbeta; ++j, c += 2) R.diffraction[R.d(i, j, k)] = polar_(line[c], line[c + 1]);
        for (int j = 0; j < R.nrad; ++j, c += 2) {
          R.added_mass[R.r(i, j, k)] = line[c];
          R.damping[R.r(i, j, k)] = line[c + 1];
        }
      }
    }
  }
  {
    Reader r(in_dir(dir, "results/FKForce.tec"));
    r.skip();
    for (int k = 0; k < R.nint; ++k) r.skip();
    for (int j = 0; j < R.nbeta; ++j) {
      r.skip();
      for (int i = 0; i < R.nw; ++i) {
        r.begin();
        (void)r.real();
        for (int k = 0; k < R.nint; ++k) {
          const double a = r.real(), ph = r.real();
          R.fk[R.d(i, j, k)] = polar_(a, ph);
        }
        r.end();
      }
    }
  }
  return R;
}

void save(const std::string& path, const std::string& text) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw FormatError("cannot open " + path + " for writing");
  const size_t n = std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  if (n != text.size()) throw FormatError("cannot write " + path);
}

// SaveTResults, Common/Results.f90:192-251
void save_results(const Results& R, const std::string& dir) {
  const double pi = 4. * std::atan(1.0);
  std::string t = "VARIABLES=\"w (rad/s)\"\n";
  for (int k = 0; k < R.nint; ++k)
    t += format_write("(A,I4,I4,A,I4,I4,A)", {"\"A", R.force_body[k], R.force_mode[k], "\" \"B", R.force_body[k],
                                                R.force_mode[k], "\""});
  for (int j = 0; j < R.nrad; ++j) {
    t += format_write("(A,I4,A,I4,A,I6,A)",
                      {"Zone t=\"Motion of body ", R.rad_body[j], " in DoF", R.rad_mode[j], "\",I=", R.nw, ",F=POINT"});
    for (int i = 0; i < R.nw; ++i) {
      std::vector<FmtItem> it{R.w[i]};
      for (int k = 0; k < R.nint; ++k) {
        it.emplace_back(R.added_mass[R.r(i, j, k)]);
        it.emplace_back(R.damping[R.r(i, j, k)]);
      }
      t += format_write("(80(X,E14.7))", it);
    }
  }
  save(in_dir(dir, "results/RadiationCoefficients.tec"), t);
  for (int excitation = 0; excitation < 2; ++excitation) {
    t = "VARIABLES=\"w (rad/s)\"\n";
    for (int k = 0; k < R.nint; ++k)
      t += format_write("(A,I4,I4,A,I4,I4,A)", {"\"abs(F", R.force_body[k], R.force_mode[k], ")\" \"angle(F",
                                                  R.force_body[k], R.force_mode[k], ")\""});
    for (int j = 0; j < R.nbeta; ++j) {
      t += format_write("(A,F7.3,A,I6,A)", {"Zone t=\"Diffraction force - beta = ", R.beta[j] * 180. / (4. * std::atan(1.0)),
                                            " deg\",I=", R.nw, ",F=POINT"});
      for (int i = 0; i < R.nw; ++i) {
        std::vector<FmtItem> it{R.w[i]};
        for (int k = 0; k < R.nint; ++k) {
          const cplx f = excitation ? R.diffraction[R.d(i, j, k)] + R.fk[R.d(i, j, k)] : R.diffraction[R.d(i, j, k)];
          it.emplace_back(abs_c(f));
          it.emplace_back(arg_c(f));
        }
        t += format_write("(80(X,E14.7))", it);
      }
    }
    save(in_dir(dir, excitation ? "results/ExcitationForce.tec" : "results/DiffractionForce.tec"), t);
  }
  // the Aquaplus-format database (Results.f90:219-250); CA/CM index the
  // force dimension with the motion index (Nradiation columns)
  for (int damping = 1; damping >= 0; --damping) {
    t = format_write("(A,I5)", {"Nb de periode : ", R.nw});
    for (int l = 0; l < R.nw; ++l) {
      t += format_write("(F7.4)", {R.w[l]});
      for (int j = 0; j < R.nrad; ++j) {
        std::vector<FmtItem> it;
        for (int k = 0; k < R.nrad; ++k) {
          if (k >= R.nint) throw FormatError("CA.dat/CM.dat need at least as many forces as motions");
          it.emplace_back(damping ? R.damping[R.r(l, j, k)] : R.added_mass[R.r(l, j, k)]);
        }
        t += format_write("(6(X,E13.6))", it);
      }
    }
    save(in_dir(dir, damping ? "results/CA.dat" : "results/CM.dat"), t);
  }
  // the header's degree sign is the Latin-1 byte of the reference source
  const std::string deg = "(\xb0)";
  t = "VARIABLES=\"Period (s)\" \"|Fx| (N/m)\" \"|Fy| (N/m)\" \"|Fz| (N/m)\" \"|Cx| (N)\" \"|Cy| (N)\" \"|Cz| (N)\" "
      "\"ang(Fx) " + deg + "\" \"ang(Fy) " + deg + "\" \"ang(Fz) " + deg + "\" \"ang(Cx) " + deg + "\" \"ang(Cy) " +
      deg + "\" \"ang(Cz) " + deg + "\"\n";
  t += format_write("(A,I2,A)", {"Zone t=\"Corps ", 1, "\""});
  for (int j = 0; j < R.nbeta; ++j) {
    t += format_write("(A,F7.3,A,I6,A)",
                      {"Zone t=\"Diffraction force - beta = ", R.beta[j] * 180. / pi, " deg\",I=", R.nw, ",F=POINT"});
    for (int l = 0; l < R.nw; ++l) {
      std::vector<FmtItem> it{R.w[l]};
      for (int k = 0; k < R.nint; ++k) it.emplace_back(abs_c(R.diffraction[R.d(l, j, k)] + R.fk[R.d(l, j, k)]));
      for (int k = 0; k < R.nint; ++k)
        it.emplace_back(180. / pi * arg_c(R.diffraction[R.d(l, j, k)] + R.fk[R.d(l, j, k)]));
      t += format_write("(F7.4,6(X,E13.6),6(X,F7.2))", it);
    }
  }
  save(in_dir(dir, "results/Fe.dat"), t);
}

// Initialize_IRF, Compute_IRF, Save_IRF (IRF.f90:66-151)
void irf(const Results& R, const std::string& dir) {
  int sw = 0;
  double dt = 0, tf = 0;
  {
    Reader r(in_dir(dir, "Nemoh.cal"));
    // the record whose first 20 characters are '--- Post processing '
    for (;;) {
      std::string l = r.peek_record().substr(0, 20);
      l.resize(20, ' ');  // CHARACTER*20, blank-padded
      if (l == "--- Post processing ") break;
      r.skip();
    }
    r.skip();
    r.begin();
    sw = (int)r.integer();
    dt = r.real();
    tf = r.real();
    r.end();
  }
  if (sw != 1) return;
  const double pi = 4. * std::atan(1.0);
  const int nt = (int)(tf / dt);
  std::vector<double> time(std::max(nt, 0));
  for (int i = 0; i < nt; ++i) time[i] = i * dt;
  const int nrad = R.nrad, nint = R.nint, nw = R.nw;
  std::vector<double> K((size_t)std::max(nt, 0) * nrad * nint), am((size_t)nrad * nint);
  auto kk = [&](int i, int j, int k) -> double& { return K[((size_t)i * nrad + j) * nint + k]; };
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nrad; ++j)
      for (int k = 0; k < nint; ++k) {
        double s = 0.;
        for (int l = 0; l + 1 < nw; ++l)
          s = s + 0.5 * (R.damping[R.r(l, j, k)] * std::cos(R.w[l] * time[i]) +
                         R.damping[R.r(l + 1, j, k)] * std::cos(R.w[l + 1] * time[i])) *
                      (R.w[l + 1] - R.w[l]);
        kk(i, j, k) = s * 2. / pi;
      }
  for (int j = 0; j < nrad; ++j)
    for (int k = 0; k < nint; ++k) {
      double a = 0.;
      for (int l = 0; l < nw; ++l) {
        double cm = 0.;
        for (int i = 0; i + 1 < nt; ++i)
          cm = cm + 0.5 * (kk(i, j, k) * std::sin(R.w[l] * time[i]) + kk(i + 1, j, k) * std::sin(R.w[l] * time[i + 1])) *
                        (time[i + 1] - time[i]);
        cm = R.added_mass[R.r(l, j, k)] + cm / R.w[l];
        a = a + cm;
      }
      am[(size_t)j * nint + k] = a / nw;
    }
  fio::Writer w;
  w.begin().text("VARIABLES=\"Time (s)\"").end();
  std::string t = w.str();
  for (int k = 0; k < nint; ++k) t += format_write("(A,I4,A,I4,A)", {"\"AddedMass ", k + 1, "\" \"IRF ", k + 1, "\""});
  for (int j = 0; j < nrad; ++j) {
    t += format_write("(A,I4,A,I6,A)", {"Zone t=\"DoF ", j + 1, "\",I=", nt, ",F=POINT"});
    for (int i = 0; i < nt; ++i) {
      std::vector<FmtItem> it{time[i]};
      for (int k = 0; k < nint; ++k) {
        it.emplace_back(am[(size_t)j * nint + k]);
        it.emplace_back(kk(i, j, k));
      }
      t += format_write("(80(X,E14.7))", it);
    }
  }
  save(in_dir(dir, "results/IRF.tec"), t);
}

}  // namespace

void postprocess(const std::string& dir) {
  const Results R = read_results(dir);
  save_results(R, dir);
  irf(R, dir);
}

}  // namespace casefiles
}  // namespace nemoh
