// Whole-case runs over the reference's files (include/nemoh_case.h): the
// preProcessor stage (preProcessor/Main.f90:26-337, Mesh.f90:93-266) and the
// Solver stage (Solver/NEMOH.f90:26-175, SOLVE_BEM.f90:86-117, OUTPUT.f90)
// with the per-frequency work on the device through the core's C ABI.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "case_io.h"
#include "nemoh_case.h"
#include "nemoh_core.h"

namespace {

using namespace nemoh::casefiles;
using nemoh::fio::FormatError;

const double kPi = 4.0 * std::atan(1.0);

struct CaseError : std::runtime_error {
  int code;
  CaseError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

std::string in_dir(const std::string& dir, const std::string& rel) {
  if (dir.empty() || dir == ".") return rel;
  return dir.back() == '/' ? dir + rel : dir + "/" + rel;
}

void ensure_dir(const std::string& p) {
  struct stat st;
  if (stat(p.c_str(), &st) == 0) return;
  if (mkdir(p.c_str(), 0755) != 0) throw CaseError(NEMOH_ERR_ARG, "cannot create directory " + p);
}

void save(const std::string& path, const std::string& text) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw CaseError(NEMOH_ERR_ARG, "cannot open " + path + " for writing");
  const size_t n = std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  if (n != text.size()) throw CaseError(NEMOH_ERR_ARG, "cannot write " + path);
}

struct Core {
  nemoh_core* h = nullptr;
  ~Core() {
    if (h) nemoh_core_destroy(h);
  }
  void check(int r, const char* what) const {
    if (r) throw CaseError(r, std::string(what) + ": " + (h ? nemoh_core_last_error(h) : "no handle"));
  }
};

// A handle for a mesh: the Solver's panel geometry (GOMG) and the wave
// tables computed here, as INITIALIZE does for the Solver.
void make_core(Core& c, int isym, const std::vector<double>& x, const std::vector<double>& y,
               const std::vector<double>& z, const std::vector<int>& p, double rho, double g, double depth) {
  const int np = (int)x.size(), nfa = (int)p.size() / 4, nsym = (isym == 1) ? 1 : 0;
  std::vector<int> m[4];
  for (int k = 0; k < 4; ++k) {
    m[k].resize(nfa);
    for (int i = 0; i < nfa; ++i) m[k][i] = p[4 * (size_t)i + k];
  }
  std::vector<double> geo[8];
  for (auto& v : geo) v.resize(nfa);
  double zer = 0;
  int r = nemoh_core_prepare_mesh(np, nfa, nsym, x.data(), y.data(), z.data(), m[0].data(), m[1].data(),
                                  m[2].data(), m[3].data(), geo[0].data(), geo[1].data(), geo[2].data(),
                                  geo[3].data(), geo[4].data(), geo[5].data(), geo[6].data(), geo[7].data(), &zer);
  if (r) throw CaseError(r, "panel geometry: a panel is degenerate or refers to a missing point");
  std::vector<double> xr(328), xz(46), t[4];
  for (auto& v : t) v.resize(328 * 46);
  r = nemoh_core_build_tables(xr.data(), xz.data(), t[0].data(), t[1].data(), t[2].data(), t[3].data());
  if (r) throw CaseError(r, "wave tables: the device is not available");
  r = nemoh_core_create(&c.h, np, nfa, nsym, x.data(), y.data(), z.data(), m[0].data(), m[1].data(), m[2].data(),
                        m[3].data(), geo[0].data(), geo[1].data(), geo[2].data(), geo[3].data(), geo[4].data(),
                        geo[5].data(), geo[6].data(), geo[7].data(), zer, g, rho, depth, xr.data(), xz.data(),
                        t[0].data(), t[1].data(), t[2].data(), t[3].data());
  if (r) throw CaseError(r, "cannot create a solver handle (is a HIP device present?)");
}

nemoh_dof to_dof(const DofCase& d) {
  nemoh_dof o{};
  o.body = d.body;
  o.kind = d.icase;
  for (int k = 0; k < 3; ++k) {
    o.dir[k] = d.dir[k];
    o.axis[k] = d.axis[k];
  }
  return o;
}

void write_mesh_files(const std::string& dir, const NemohCal& cal, const PreMesh& m) {
  ensure_dir(in_dir(dir, "mesh"));
  ensure_dir(in_dir(dir, "results"));
  save(in_dir(dir, "mesh/L12.dat"), l12_text(m));
  save(in_dir(dir, "mesh/L10.dat"), l10_text(m));
  save(in_dir(dir, "mesh/Mesh.tec"), mesh_tec_text(m));
  save(in_dir(dir, "mesh/Freesurface.dat"), freesurface_dat_text(cal));
  save(in_dir(dir, "mesh/Kochin.dat"), kochin_dat_text(cal));
  save(in_dir(dir, "results/index.dat"), index_dat_text(cal));
}

// preProcessor/Main.f90:186-261: force weights, body conditions and FK
// forces, from the device, on the mesh files' coordinates.
void pre_stage(const std::string& dir) {
  const NemohCal cal = read_nemoh_cal(in_dir(dir, "Nemoh.cal"));
  const PreMesh m = read_body_meshes(cal, in_dir(dir, "Nemoh.cal"), dir);
  write_mesh_files(dir, cal, m);
  const int nfa = m.nfa(), ntot = nfa << ((m.isym == 1) ? 1 : 0);
  const int nrad = (int)cal.rad.size(), nint = (int)cal.force.size(), nbeta = cal.nbeta, nw = cal.nw;
  const int per = nbeta + nrad;
  Core c;
  make_core(c, m.isym, m.x, m.y, m.z, m.p, cal.rho, cal.g, cal.depth);
  std::vector<nemoh_dof> force(nint), rad(nrad);
  for (int k = 0; k < nint; ++k) force[k] = to_dof(cal.force[k]);
  for (int k = 0; k < nrad; ++k) rad[k] = to_dof(cal.rad[k]);
  std::vector<cplx> nvel((size_t)per * nw * ntot), fk((size_t)nw * nbeta * nint);
  std::vector<double> fnds((size_t)nint * ntot, 0.0);
  std::vector<double> buf, pinc, nds_cm;
  // diffraction problems and incident pressures, up to 16 headings a call
  c.check(nemoh_core_set_dofs(c.h, m.cpanel.data(), 0, nullptr, nint, force.data(), cal.xeff, cal.yeff),
          "set_dofs");
  bool have_nds = false;
  for (int i = 0; i < nw; ++i)
    for (int j0 = 0; j0 < nbeta; j0 += 16) {
      const int nb = std::min(16, nbeta - j0);
      buf.assign(2 * (size_t)ntot * nb, 0.0);
      pinc.assign(2 * (size_t)ntot * nb, 0.0);
      nds_cm.assign(std::max<size_t>(1, (size_t)nint * ntot), 0.0);
      c.check(nemoh_core_body_conditions(c.h, cal.w[i], nb, cal.beta.data() + j0, buf.data(), pinc.data(),
                                         have_nds ? nullptr : nds_cm.data()),
              "body conditions");
      if (!have_nds) {
        for (int k = 0; k < nint; ++k)
          for (int cc = 0; cc < ntot; ++cc) fnds[(size_t)k * ntot + cc] = nds_cm[k + (size_t)nint * cc];
        have_nds = true;
      }
      for (int q = 0; q < nb; ++q) {
        const size_t col = (size_t)i * per + j0 + q;
        for (int cc = 0; cc < ntot; ++cc)
          nvel[col * ntot + cc] = cplx(buf[2 * ((size_t)q * ntot + cc)], buf[2 * ((size_t)q * ntot + cc) + 1]);
        for (int k = 0; k < nint; ++k) {  // FKForce = FKForce - PRESSURE(c) FNDS(k,c), Main.f90:227-232
          cplx f(0.0, 0.0);
          for (int cc = 0; cc < ntot; ++cc) {
            const double r = fnds[(size_t)k * ntot + cc];
            const cplx p(pinc[2 * ((size_t)q * ntot + cc)], pinc[2 * ((size_t)q * ntot + cc) + 1]);
            f -= cplx(p.real() * r, p.imag() * r);
          }
          fk[((size_t)i * nbeta + j0 + q) * nint + k] = f;
        }
      }
    }
  // radiation problems (the same at every frequency), up to 16 a call
  for (int k0 = 0; k0 < nrad; k0 += 16) {
    const int nr = std::min(16, nrad - k0);
    c.check(nemoh_core_set_dofs(c.h, m.cpanel.data(), nr, rad.data() + k0, nint, force.data(), cal.xeff, cal.yeff),
            "set_dofs");
    buf.assign(2 * (size_t)ntot * nr, 0.0);
    nds_cm.assign(std::max<size_t>(1, (size_t)nint * ntot), 0.0);
    c.check(nemoh_core_body_conditions(c.h, cal.w[0], 0, nullptr, buf.data(), nullptr, nds_cm.data()),
            "body conditions");
    if (!have_nds) {
      for (int k = 0; k < nint; ++k)
        for (int cc = 0; cc < ntot; ++cc) fnds[(size_t)k * ntot + cc] = nds_cm[k + (size_t)nint * cc];
      have_nds = true;
    }
    for (int i = 0; i < nw; ++i)
      for (int q = 0; q < nr; ++q) {
        const size_t col = (size_t)i * per + nbeta + k0 + q;
        for (int cc = 0; cc < ntot; ++cc)
          nvel[col * ntot + cc] = cplx(buf[2 * ((size_t)q * ntot + cc)], buf[2 * ((size_t)q * ntot + cc) + 1]);
      }
  }
  save(in_dir(dir, "mesh/Integration.dat"), integration_text(nint, ntot, fnds));
  save(in_dir(dir, "Normalvelocities.dat"),
       normal_velocities_text(cal, ntot, nvel, cal.nx > 0 ? 1 : 0, cal.ntheta > 0 ? 1 : 0));
  save(in_dir(dir, "results/FKForce.tec"), fkforce_tec_text(cal, fk));
  save(in_dir(dir, "results/FKForce.dat"), fkforce_dat_text(cal, fk));
}

// Solver/NEMOH.f90:52-165 with SOLVE_BVP's outputs (SOLVE_BEM.f90:86-117).
void solve_stage(const std::string& dir, int precision) {
  const Environment env = read_environment(in_dir(dir, "Nemoh.cal"));
  const SolverMesh m = read_solver_mesh(dir);
  const int nsym = (m.isym == 1) ? 1 : 0, nfa = m.nfa(), ntot = nfa << nsym;
  const BodyConditions bc = read_body_conditions(in_dir(dir, "Normalvelocities.dat"), ntot);
  int nint = 0;
  const std::vector<double> nds_rm = read_integration(in_dir(dir, "mesh/Integration.dat"), ntot, &nint);
  const std::vector<double> theta = read_kochin_dat(in_dir(dir, "mesh/Kochin.dat"));
  const FreeSurfaceMesh fs = read_freesurface_dat(in_dir(dir, "mesh/Freesurface.dat"));
  ensure_dir(in_dir(dir, "results"));
  Core c;
  make_core(c, m.isym, m.x, m.y, m.z, m.p, env.rho, env.g, env.depth);
  c.check(nemoh_core_set_precision(c.h, precision, 0.0, 0), "precision");
  const double xeff = env.xeff, yeff = nsym ? 0.0 : env.yeff;  // INITIALIZATION.f90:54-58
  std::vector<double> nds_cm((size_t)nint * ntot);
  for (int k = 0; k < nint; ++k)
    for (int cc = 0; cc < ntot; ++cc) nds_cm[k + (size_t)nint * cc] = nds_rm[(size_t)k * ntot + cc];
  std::vector<cplx> force((size_t)nint * bc.nprob);
  std::vector<double> fsxyz;
  for (int i = 0; i < fs.npoints; ++i) {
    fsxyz.push_back(fs.x[i]);
    fsxyz.push_back(fs.y[i]);
    fsxyz.push_back(0.0);
  }
  int j0 = 0;
  while (j0 < bc.nprob) {
    int m_ = 1;  // consecutive problems at the same frequency, up to 16
    while (j0 + m_ < bc.nprob && m_ < 16 && bc.omega[j0 + m_] == bc.omega[j0]) ++m_;
    const double period = 2. * kPi / bc.omega[j0];  // NEMOH.f90:135
    std::vector<double> nv(2 * (size_t)ntot * m_), pr(2 * (size_t)ntot * m_), fo(2 * (size_t)std::max(1, nint) * m_);
    for (int q = 0; q < m_; ++q)
      for (int cc = 0; cc < ntot; ++cc) {
        const cplx v = bc.nvel[(size_t)(j0 + q) * ntot + cc];
        nv[2 * ((size_t)q * ntot + cc)] = v.real();
        nv[2 * ((size_t)q * ntot + cc) + 1] = v.imag();
      }
    c.check(nemoh_core_solve_frequency(c.h, period, m_, nv.data(), pr.data(), nint,
                                       nint > 0 ? nds_cm.data() : nullptr, nint > 0 ? fo.data() : nullptr),
            "solve");
    for (int q = 0; q < m_; ++q)
      for (int k = 0; k < nint; ++k)
        force[(size_t)(j0 + q) * nint + k] = cplx(fo[2 * (k + (size_t)nint * q)], fo[2 * (k + (size_t)nint * q) + 1]);
    const double w = 2. * kPi / period;  // SOLVE_BEM.f90:63
    bool want_k = false, want_fs = false;
    for (int q = 0; q < m_; ++q) {
      want_k |= bc.sw_kochin[j0 + q] == 1;
      want_fs |= bc.sw_fs[j0 + q] == 1;
    }
    std::vector<double> hk, phi;
    if (want_k && !theta.empty()) {
      hk.assign(2 * theta.size() * m_, 0.0);
      c.check(nemoh_core_kochin(c.h, xeff, yeff, (int)theta.size(), theta.data(), hk.data()), "Kochin function");
    }
    if (want_fs && fs.npoints > 0) {
      phi.assign(2 * (size_t)fs.npoints * m_, 0.0);
      c.check(nemoh_core_potential(c.h, fs.npoints, fsxyz.data(), phi.data()), "free-surface potential");
    }
    for (int q = 0; q < m_; ++q) {
      const int j = j0 + q;
      const std::string tag = problem_tag(j + 1);
      if (bc.sw_kochin[j] == 1) {
        std::vector<cplx> h(theta.size());
        for (size_t t = 0; t < theta.size(); ++t)
          h[t] = cplx(hk[2 * (t + theta.size() * q)], hk[2 * (t + theta.size() * q) + 1]);
        save(in_dir(dir, "results/Kochin." + tag + ".dat"), kochin_text(theta, h.data()));
      }
      if (bc.sw_fs[j] == 1 && fs.npoints > 0) {
        const double a = w / env.g;  // ETA = II*W/G*PHI
        std::vector<cplx> eta(fs.npoints);
        for (int p = 0; p < fs.npoints; ++p) {
          const double pr_ = phi[2 * (p + (size_t)fs.npoints * q)], pi_ = phi[2 * (p + (size_t)fs.npoints * q) + 1];
          eta[p] = cplx(0.0 * pr_ - a * pi_, 0.0 * pi_ + a * pr_);
        }
        save(in_dir(dir, "results/freesurface." + tag + ".dat"), freesurface_text(fs, eta.data()));
      }
      if (bc.sw_potential[j] == 1) {
        std::vector<cplx> p(ntot);
        for (int cc = 0; cc < ntot; ++cc) p[cc] = cplx(pr[2 * ((size_t)q * ntot + cc)], pr[2 * ((size_t)q * ntot + cc) + 1]);
        save(in_dir(dir, "results/pressure." + tag + ".dat"), pressure_text(m, p.data()));
      }
    }
    j0 += m_;
  }
  save(in_dir(dir, "results/Forces.dat"), forces_text(bc, nint, force));
}

int report(int code, const std::string& what, char* msg, int msglen) {
  if (msg && msglen > 0) std::snprintf(msg, (size_t)msglen, "%s", what.c_str());
  return code;
}

}  // namespace

extern "C" {

int nemoh_case_run(const char* case_dir, int stages, int precision, char* msg, int msglen) {
  if (msg && msglen > 0) msg[0] = '\0';
  if (!case_dir || !(stages & (NEMOH_STAGE_PRE | NEMOH_STAGE_SOLVE | NEMOH_STAGE_POST)) ||
      (precision != NEMOH_FP64 && precision != NEMOH_MIXED))
    return report(NEMOH_ERR_ARG, "bad arguments", msg, msglen);
  try {
    if (stages & NEMOH_STAGE_PRE) pre_stage(case_dir);
    if (stages & NEMOH_STAGE_SOLVE) solve_stage(case_dir, precision);
    if (stages & NEMOH_STAGE_POST) postprocess(case_dir);
  } catch (const CaseError& e) {
    return report(e.code, e.what(), msg, msglen);
  } catch (const FormatError& e) {
    return report(NEMOH_ERR_ARG, e.what(), msg, msglen);
  } catch (const std::exception& e) {
    return report(NEMOH_ERR_ARG, e.what(), msg, msglen);
  }
  return NEMOH_OK;
}

int nemoh_case_write_mesh_files(const char* case_dir, char* msg, int msglen) {
  if (msg && msglen > 0) msg[0] = '\0';
  if (!case_dir) return report(NEMOH_ERR_ARG, "bad arguments", msg, msglen);
  try {
    const std::string dir = case_dir;
    const NemohCal cal = read_nemoh_cal(in_dir(dir, "Nemoh.cal"));
    const PreMesh m = read_body_meshes(cal, in_dir(dir, "Nemoh.cal"), dir);
    write_mesh_files(dir, cal, m);
  } catch (const std::exception& e) {
    return report(NEMOH_ERR_ARG, e.what(), msg, msglen);
  }
  return NEMOH_OK;
}

int nemoh_format_list(int n, const double* v, char* out, size_t cap) {
  if (n < 0 || (n > 0 && !v) || !out || cap == 0) return -1;
  nemoh::fio::Writer w;
  w.begin();
  for (int i = 0; i < n; ++i) w.real(v[i]);
  w.end();
  const std::string& s = w.str();
  if (s.size() + 1 > cap) return -1;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size();
}

int nemoh_format_e(double v, int w, int d, char* out, size_t cap) {
  if (w <= 0 || d <= 0 || !out || cap == 0) return -1;
  const std::string s = nemoh::fio::edit_E(v, w, d);
  if (s.size() + 1 > cap) return -1;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size();
}

}  // extern "C"
