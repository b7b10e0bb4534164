"""Generate the golden fixtures from the FP64 build of the reference.

Runs only where the reference sources are present (it builds them with
oracle/Makefile.ref into oracle/_ref/).  It

* runs the reference preProcessor + Solver on the four reference
  Verification cases (Cylinder, NonSymmetrical, 2Bodies, Postprocessing),
  on the infinite-depth 512-panel cylinder (R=10 m, T=5 m, 20 frequencies
  0.1-2.0 rad/s, half mesh) and on a small finite-depth barge, and stores
  the Solver's inputs (mesh/L12.dat geometry, Normalvelocities.dat,
  Integration.dat) with its output (results/Forces.dat) as case_*.npz;
  the same for one body of each shape of the batched workload (c4cyl,
  c4box, c4sph) and for the full-size 4000-panel barge at three
  frequencies (barge4000, whose wall time goes to profiles/);
* runs oracle/ref_harness.f90 (linked against the reference modules) to
  dump GG samples, the CREK tables, CINT rules, X0, LISC fits and every
  pair of VAV*+VNS* on two small meshes, stored as fn_*.npz / pairs_*.npz.

Edits to the reference inputs (all outside the Forces.dat path): Kochin
and free-surface post-processing switched off; Cylinder's Ntheta literal
`0.` made the integer `0` (preProcessor/Main.f90:138 reads an integer);
2Bodies reduced from 40 to 4 frequencies to keep the fixture small.

Usage: python tests/golden/make_golden.py [--ref /root/reference]
"""
from __future__ import annotations

import argparse
import re
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from nemoh_amd import io as nio  # noqa: E402
from nemoh_amd import mesh as nmesh  # noqa: E402

OUT = Path(__file__).resolve().parent
REF_BIN = ROOT / "oracle" / "_ref"


def build_ref(ref: str) -> None:
    subprocess.run(["make", "-s", "-f", "oracle/Makefile.ref", f"REF={ref}"], cwd=ROOT, check=True)


def run_case(work: Path, name: str, nbodies_ntot=None) -> None:
    (work / "mesh").mkdir(exist_ok=True)
    (work / "results").mkdir(exist_ok=True)
    (work / "ID.dat").write_text(f"1\n.\n")
    if not (work / "input.txt").exists():
        (work / "input.txt").write_text("--- Calculation parameters ---\n0\t! Indiq_solver\n")
    for exe in ("preproc", "solver"):
        r = subprocess.run([str(REF_BIN / exe)], cwd=work, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: {exe} failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")


def collect(work: Path, name: str, rho, g, depth) -> None:
    m = nmesh.read_dat(work / "mesh" / "L12.dat")
    ntot = m.npanels * 2 ** m.isym
    bc = nio.read_normal_velocities(work / "Normalvelocities.dat", ntot)
    nds = nio.read_integration(work / "mesh" / "Integration.dat", ntot)
    table = nio.read_forces_dat(work / "results" / "Forces.dat", nds.shape[0], bc.omega.size)
    np.savez_compressed(OUT / f"case_{name}.npz", vertices=m.vertices, panels=m.panels,
                        isym=m.isym, rho=rho, g=g, depth=depth, omega=bc.omega, kind=bc.kind,
                        nvel=bc.nvel, nds=nds, forces_table=table)
    print(f"case {name}: N={m.npanels} isym={m.isym} problems={bc.omega.size} nint={nds.shape[0]}")


def edit_cal(text: str, repl: dict) -> str:
    lines = text.split("\n")
    for k, v in repl.items():
        lines[k - 1] = v
    return "\n".join(lines)


def verification_cases(ref: Path, tmp: Path) -> None:
    spec = {
        # name: (line edits 1-based, depth)
        "Cylinder": ({32: "0\t0.\t180.\t! Kochin", 33: "0\t10\t100.\t100.\t! FS"}, 0.0),
        "NonSymmetrical": ({32: "0\t0.\t180.\t! Kochin", 33: "0\t2\t1000.\t2.\t! FS"}, 0.0),
        "2Bodies": ({45: "4\t0.1\t2.0\t! Number of wave frequencies",
                     50: "0\t0.\t180.\t! Kochin", 51: "0\t2\t1000.\t2.\t! FS"}, 20.0),
        "Postprocessing": ({32: "0\t0.\t180.\t! Kochin", 33: "0\t50\t400.\t400.\t! FS"}, 10.0),
    }
    for name, (edits, depth) in spec.items():
        src = ref / "Verification" / name
        work = tmp / name
        work.mkdir()
        for f in src.iterdir():
            if f.is_file() and f.suffix in (".dat", ".cal", ".txt"):
                shutil.copy(f, work / f.name)
        cal = (work / "Nemoh.cal").read_text(errors="replace")
        (work / "Nemoh.cal").write_text(edit_cal(cal, edits))
        run_case(work, name)
        collect(work, name, 1000.0, 9.81, depth)


def synthetic_cases(tmp: Path) -> None:
    # infinite-depth cylinder R=10 T=5, half mesh 512 panels, 20 omegas
    cyl = nmesh.cylinder(10.0, 5.0, 32, 8, 8, half=True)
    assert cyl.npanels == 512
    work = tmp / "cyl512"
    work.mkdir()
    cyl.write_dat(work / "cyl.dat")
    body = nio.BodySpec("cyl.dat", cyl.npoints, cyl.npanels, nio.six_dof(), nio.six_dof())
    nio.write_nemoh_cal(work / "Nemoh.cal", 1025.0, 9.81, 0.0, [body], 20, 0.1, 2.0, 1, 0.0, 0.0)
    run_case(work, "cyl512")
    collect(work, "cyl512", 1025.0, 9.81, 0.0)
    # finite-depth barge 100x20x10 at h=50, no symmetry, 3 headings;
    # omega = 2.0 takes the infinite-depth path after finite-depth ones
    barge = nmesh.box(100.0, 20.0, 10.0, 18, 4, 2)
    work = tmp / "barge_fd"
    work.mkdir()
    barge.write_dat(work / "barge.dat")
    body = nio.BodySpec("barge.dat", barge.npoints, barge.npanels, nio.six_dof(), nio.six_dof())
    nio.write_nemoh_cal(work / "Nemoh.cal", 1025.0, 9.81, 50.0, [body], 3, 0.2, 2.0, 3, 0.0, 90.0)
    run_case(work, "barge_fd")
    collect(work, "barge_fd", 1025.0, 9.81, 50.0)


def barge4000_case(tmp: Path) -> None:
    """The full-size finite-depth barge (100 x 20 x 10 m, 4000 panels, no
    symmetry, h = 50 m, 6 DOF + headings 0/45/90) at three frequencies of its
    0.2-2.0 rad/s band: 0.2 and 1.1 take the finite-depth path, 2.0 the
    infinite-depth one (k h >= 20) with the NQ the finite-depth calls left.
    The reference's wall time is written to profiles/."""
    import time
    barge = nmesh.box(100.0, 20.0, 10.0, 90, 20, 10)
    assert barge.npanels == 4000
    work = tmp / "barge4000"
    work.mkdir()
    barge.write_dat(work / "barge.dat")
    body = nio.BodySpec("barge.dat", barge.npoints, barge.npanels, nio.six_dof(), nio.six_dof())
    nio.write_nemoh_cal(work / "Nemoh.cal", 1025.0, 9.81, 50.0, [body], 3, 0.2, 2.0, 3, 0.0, 90.0)
    t0 = time.perf_counter()
    run_case(work, "barge4000")
    wall = time.perf_counter() - t0
    collect(work, "barge4000", 1025.0, 9.81, 50.0)
    (ROOT / "profiles" / "cpu_reference_barge4000.txt").write_text(
        f"reference preProcessor + Solver, flang -O2 FP64 build of the unmodified sources (oracle/Makefile.ref), "
        f"one process on one core: barge 100x20x10 m, 4000 panels, h = 50 m, 3 frequencies (0.2, 1.1, 2.0 rad/s) "
        f"x 9 problems: {wall:.1f} s wall, {wall / 3:.1f} s per frequency\n"
        f"command: python tests/golden/make_golden.py --only barge4000\n")
    print(f"barge4000: {wall:.1f} s")


def c4_bodies_cases(tmp: Path) -> None:
    """One body of each shape of the batched workload (cylinder, box, prolate
    spheroid), drawn by nemoh_amd.serving.random_body_parts from the first
    seeds that give that shape with 300-600 panels: 6 DOF + heading 0, the
    8 frequencies 0.2-2.0 rad/s, infinite depth."""
    from nemoh_amd import serving
    found = {}
    seed = 0
    while len(found) < 3:
        rng = np.random.default_rng(seed)
        kind = int(np.random.default_rng(seed).integers(0, 3))
        m, _, omegas, _ = serving.random_body_parts(rng, n_range=(300, 600))
        if kind not in found and 300 <= m.npanels <= 600:
            found[kind] = (seed, m)
        seed += 1
    for kind, name in ((0, "c4cyl"), (1, "c4box"), (2, "c4sph")):
        seed, m = found[kind]
        work = tmp / name
        work.mkdir()
        m.write_dat(work / "body.dat")
        body = nio.BodySpec("body.dat", m.npoints, m.npanels, nio.six_dof(), nio.six_dof())
        nio.write_nemoh_cal(work / "Nemoh.cal", 1025.0, 9.81, 0.0, [body], 8, 0.2, 2.0, 1, 0.0, 0.0)
        run_case(work, name)
        collect(work, name, 1025.0, 9.81, 0.0)
        print(f"{name}: seed {seed}, {m.npanels} panels")


def harness(cmds: str) -> list:
    r = subprocess.run([str(REF_BIN / "harness")], input=cmds, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    return [ln for ln in r.stdout.split("\n") if ln.strip()]


def fn_fixtures() -> None:
    rng = np.random.default_rng(20140101)
    # GG over its four regimes and the branch boundaries
    z = np.concatenate([
        rng.uniform(-40, -16, 200) + 1j * rng.uniform(-30, 30, 200),    # Re z <= -16
        rng.uniform(-16, 10, 200) + 1j * rng.choice([-1, 1], 200) * rng.uniform(10, 40, 200),
        rng.uniform(-0.5, 8, 300) + 1j * rng.uniform(-10, 10, 300),      # log series
        rng.uniform(-16, -0.5, 300) + 1j * rng.uniform(-10, 10, 300),    # rational
        np.array([-16 + 0j, -0.5 + 1j, -3 + 10j, -3 - 10j, 0.5 + 0j, -2 + 0j, -25 + 2j]),
    ])
    lines = harness("gg\n%d\n" % z.size + "".join(f"{v.real:.17e} {v.imag:.17e}\n" for v in z) + "end\n")
    a = np.array([[float(t) for t in ln.split()] for ln in lines])
    np.savez_compressed(OUT / "fn_gg.npz", z=a[:, 0] + 1j * a[:, 1], gg=a[:, 2] + 1j * a[:, 3])

    lines = harness("crek\nend\n")
    vals = np.array([[float(t) for t in ln.split()] for ln in lines[328 + 46:]])
    xr = np.array([float(ln) for ln in lines[:328]])
    xz = np.array([float(ln) for ln in lines[328:374]])
    # rows are (i fastest, j slowest): reshape to (46, 328, 4)
    v = vals.reshape(46, 328, 4)
    np.savez_compressed(OUT / "fn_crek.npz", xr=xr, xz=xz, apd1x=v[:, :, 0].T, apd1z=v[:, :, 1].T,
                        apd2x=v[:, :, 2].T, apd2z=v[:, :, 3].T)

    akk = np.array([0.05, 0.4, 0.41, 2.5, 2.6, 4.0, 5.0, 8.0, 9.0, 25.0, 26.0, 40.0, 41.0, 100.0])
    lines = harness("cinti\n%d\n" % akk.size + "".join(f"{v:.17e}\n" for v in akk) + "end\n")
    i, nq, nqmod = 0, [], []
    qq, cq = np.zeros((akk.size, 101)), np.zeros((akk.size, 101))
    for k in range(akk.size):
        t = lines[i].split()
        n = int(t[1])
        nq.append(n)
        nqmod.append(int(t[2]))
        r = np.array([[float(v) for v in lines[i + 1 + q].split()] for q in range(n)])
        qq[k, :n], cq[k, :n] = r[:, 0], r[:, 1]
        i += 1 + n
    np.savez_compressed(OUT / "fn_cint.npz", akk=akk, nq=np.array(nq), nq_module=np.array(nqmod), qq=qq, cq=cq)

    ak = np.concatenate([np.logspace(-3, np.log10(19.9), 40), [0.0408, 0.25, 1.0, 5.0]])
    lines = harness("x0\n%d\n" % ak.size + "".join(f"{v:.17e}\n" for v in ak) + "end\n")
    a = np.array([[float(t) for t in ln.split()] for ln in lines])
    np.savez_compressed(OUT / "fn_x0.npz", ak=a[:, 0], x0=a[:, 1])

    pairs = []
    for h in (10.0, 20.0, 50.0):
        for w in np.linspace(0.2, 2.5, 12):
            akh0 = w * w * h / 9.81
            if akh0 >= 20:
                continue
            amh = float(harness(f"x0\n1\n{akh0:.17e}\nend\n")[0].split()[1])
            pairs.append((amh * np.tanh(amh), amh))
    pairs.append((0.5, 0.8))
    # the solver's own call sequence (AKH = w^2 h / g, AMH = X0(AKH)) over a
    # 40-frequency sweep 0.2-2.0 rad/s at h = 50 m, which includes fits where
    # an intermediate step keeps no exponent
    for w in np.linspace(0.2, 2.0, 40):
        akh0 = w * w * 50.0 / 9.81
        if akh0 < 20:
            pairs.append((akh0, float(harness(f"x0\n1\n{akh0:.17e}\nend\n")[0].split()[1])))
    lines = harness("lisc\n%d\n" % len(pairs) + "".join(f"{p[0]:.17e} {p[1]:.17e}\n" for p in pairs) + "end\n")
    i, nexp, lam, ar, akh, amh = 0, [], np.zeros((len(pairs), 31)), np.zeros((len(pairs), 31)), [], []
    for k in range(len(pairs)):
        t = lines[i].split()
        akh.append(float(t[0]))
        amh.append(float(t[1]))
        n = int(t[2])
        i += 1
        r = np.array([[float(v) for v in lines[i + q].split()] for q in range(n)])
        i += n
        nexp.append(n)
        lam[k, :n], ar[k, :n] = r[:, 0], r[:, 1]
    np.savez_compressed(OUT / "fn_lisc.npz", akh=np.array(akh), amh=np.array(amh),
                        nexp=np.array(nexp), ambda=lam, ar=ar)

    # FF on the fit interval for three (AKH, AMH) of the solver's sweep
    pts = [(x, pairs[k][0], pairs[k][1]) for k in (-40, -20, -3) for x in np.linspace(-0.1, 20.0, 24)]
    lines = harness("ff\n%d\n" % len(pts) + "".join(f"{x:.17e} {a_:.17e} {m_:.17e}\n" for x, a_, m_ in pts) + "end\n")
    np.savez_compressed(OUT / "fn_ff.npz", pts=np.array(pts), ff=np.array([float(ln) for ln in lines]))

    # HOUSRS on random least-squares systems up to LISC's largest (91 x 30)
    shapes, mats, sols = [], [], []
    for nl, ncc in ((6, 2), (20, 7), (57, 16), (91, 30)):
        a_ = rng.standard_normal((nl, ncc + 1))
        lines = harness(f"housrs\n{nl} {ncc}\n" + "".join(" ".join(f"{v:.17e}" for v in row) + "\n" for row in a_)
                        + "end\n")
        shapes.append((nl, ncc))
        mats.append(np.pad(a_, ((0, 91 - nl), (0, 31 - ncc - 1))))
        sols.append(np.pad(np.array([float(ln) for ln in lines]), (0, 30 - ncc)))
    np.savez_compressed(OUT / "fn_housrs.npz", shapes=np.array(shapes), a=np.array(mats), x=np.array(sols))

    # SPRBM on polynomials with known real roots and complex pairs, c[0] the
    # constant term (as LISC builds them, leading coefficient last)
    polys = []
    for deg in (4, 7, 12, 20):
        nc = deg // 3
        roots = list(rng.uniform(0.05, 0.95, deg - 2 * nc))
        for _ in range(nc):
            rho, th = rng.uniform(0.2, 0.9), rng.uniform(0.3, 2.8)
            roots += [rho * np.exp(1j * th), rho * np.exp(-1j * th)]
        polys.append(np.real(np.poly(roots))[::-1])
    out = {"ier": [], "ir": [], "c": [], "rr": [], "ri": []}
    for c in polys:
        lines = harness(f"sprbm\n{c.size}\n" + " ".join(f"{v:.17e}" for v in c) + "\nend\n")
        ir, ier = map(int, lines[0].split())
        r = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + c.size]])
        out["ier"].append(ier)
        out["ir"].append(ir)
        out["c"].append(np.pad(c, (0, 21 - c.size)))
        out["rr"].append(np.pad(r[:, 0], (0, 21 - c.size)))
        out["ri"].append(np.pad(r[:, 1], (0, 21 - c.size)))
    np.savez_compressed(OUT / "fn_sprbm.npz", deg=np.array([p_.size - 1 for p_ in polys]),
                        **{k: np.array(v) for k, v in out.items()})
    print("function fixtures written")


def mesh_cmd(m: nmesh.Mesh, depth: float) -> str:
    s = f"mesh\n{m.npoints} {m.npanels} {m.isym} {depth:.17e}\n"
    s += "".join(f"{x:.17e} {y:.17e} {z:.17e}\n" for x, y, z in m.vertices)
    s += "".join(f"{a} {b} {c} {d}\n" for a, b, c, d in m.panels + 1)
    return s


def parse_pairs(lines, i, nfa):
    head = lines[i].split()
    i += 1
    info = {"path": head[0]}
    if head[0] == "FD":
        info["amh"], info["nq"], info["nexp"] = float(head[1]), int(head[2]), int(head[3])
        r = np.array([[float(v) for v in lines[i + q].split()] for q in range(info["nexp"])])
        info["ambda"], info["ar"] = r[:, 0], r[:, 1]
        i += info["nexp"]
    else:
        info["am0"], info["nq"] = float(head[1]), int(head[2])
    out = np.zeros((nfa, nfa, 16))
    for p in range(nfa * nfa):
        a = lines[i].split()
        isp, ifp = int(a[0]) - 1, int(a[1]) - 1
        row = [float(v) for v in a[2:]] + [float(v) for v in lines[i + 1].split()] + \
              [float(v) for v in lines[i + 2].split()]
        out[isp, ifp] = row
        i += 3
    return i, info, out


def lid_cylinder() -> nmesh.Mesh:
    """Small half cylinder plus a ring of free-surface lid panels (z = 0),
    whose collocation points sit above ZER and take the near-free-surface
    branch of the wave Green function."""
    c = nmesh.cylinder(5.0, 4.0, 6, 3, 2, half=True)
    pool_v = [list(v) for v in c.vertices]
    quads = [list(p) for p in c.panels]

    def vid(p):
        for k, v in enumerate(pool_v):
            if np.allclose(v, p, atol=1e-9):
                return k
        pool_v.append(list(p))
        return len(pool_v) - 1
    th = np.linspace(0, np.pi, 7)
    for a in range(6):
        r0, r1 = 2.5, 5.0
        q = [vid((r0 * np.cos(th[a]), r0 * np.sin(th[a]), 0.0)),
             vid((r0 * np.cos(th[a + 1]), r0 * np.sin(th[a + 1]), 0.0)),
             vid((r1 * np.cos(th[a + 1]), r1 * np.sin(th[a + 1]), 0.0)),
             vid((r1 * np.cos(th[a]), r1 * np.sin(th[a]), 0.0))]
        quads.append(q)
    return nmesh.Mesh(np.array(pool_v), np.array(quads, dtype=np.int32), 1)


def pair_fixtures() -> None:
    # symmetric mesh with lid panels: FD pairs first, then INFD pairs (which
    # inherit NQ from the FD call), then INFD at a high frequency
    m = lid_cylinder()
    cmds = mesh_cmd(m, 30.0) + "pairs\n6.0\ndepth\n0.0\npairs\n6.0\npairs\n1.2\nend\n"
    lines = harness(cmds)
    zer = float(lines[0])
    geo = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + m.npanels]])
    i = 1 + m.npanels
    i, info_fd, fd = parse_pairs(lines, i, m.npanels)
    i, info_i1, i1 = parse_pairs(lines, i, m.npanels)
    i, info_i2, i2 = parse_pairs(lines, i, m.npanels)
    np.savez_compressed(OUT / "pairs_lidcyl.npz", vertices=m.vertices, panels=m.panels, isym=1,
                        zer=zer, geo=geo, depth_fd=30.0, period_fd=6.0, fd=fd,
                        fd_amh=info_fd["amh"], fd_nq=info_fd["nq"], fd_ambda=info_fd["ambda"],
                        fd_ar=info_fd["ar"], period_i1=6.0, infd1=i1, infd1_nq=info_i1["nq"],
                        period_i2=1.2, infd2=i2, infd2_nq=info_i2["nq"])
    # non-symmetric barge, finite depth then infinite depth
    b = nmesh.box(20.0, 8.0, 4.0, 5, 2, 2)
    cmds = mesh_cmd(b, 12.0) + "pairs\n5.0\ndepth\n0.0\npairs\n2.5\nend\n"
    lines = harness(cmds)
    zer = float(lines[0])
    geo = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + b.npanels]])
    i = 1 + b.npanels
    i, info_fd, fd = parse_pairs(lines, i, b.npanels)
    i, info_i1, i1 = parse_pairs(lines, i, b.npanels)
    np.savez_compressed(OUT / "pairs_barge.npz", vertices=b.vertices, panels=b.panels, isym=0,
                        zer=zer, geo=geo, depth_fd=12.0, period_fd=5.0, fd=fd,
                        fd_amh=info_fd["amh"], fd_nq=info_fd["nq"], fd_ambda=info_fd["ambda"],
                        fd_ar=info_fd["ar"], period_i1=2.5, infd1=i1, infd1_nq=info_i1["nq"])
    print("pair fixtures written")


def wavenumber_fixture() -> None:
    """The preProcessor's own wave number (Common/Environment.f90:45-84) over
    frequencies and depths that cover its deep-water branch (k h > 20 or
    h = 0) and the bisection."""
    rows = [(w, h, 9.81) for h in (0.0, 2.0, 10.0, 50.0, 200.0, 1000.0) for w in np.linspace(0.05, 3.0, 40)]
    rows += [(0.8, 50.0, 9.80665), (1.3, 12.5, 9.81), (2.0, 50.0, 9.81)]
    lines = harness("wavenumber\n%d\n" % len(rows) + "".join(f"{a:.17e} {b:.17e} {c:.17e}\n" for a, b, c in rows)
                    + "end\n")
    a = np.array([[float(v) for v in ln.split()] for ln in lines])
    np.savez_compressed(OUT / "fn_wavenumber.npz", w=a[:, 0], depth=a[:, 1], g=a[:, 2], k=a[:, 3])
    print(f"wavenumber: {a.shape[0]} cases")


def kochin_fixture() -> None:
    """COMPUTE_KOCHIN (COMPUTE_KOCHIN.f90:25-153) of the reference build on
    seeded random sources: a half cylinder (y-symmetry) and a barge, in
    infinite depth and at h = 50 m on both sides of its k h = 18 switch, with
    an off-origin wave reference point."""
    rng = np.random.default_rng(2718)
    cases = []
    for name, m in (("cyl", nmesh.cylinder(10.0, 5.0, 16, 4, 4, half=True)),
                    ("barge", nmesh.box(100.0, 20.0, 10.0, 18, 4, 2))):
        for depth, kws in ((0.0, (0.05, 0.6)), (50.0, (0.05, 0.3, 0.5))):
            for kw in kws:
                thetas = np.linspace(0.0, 2.0 * np.pi, 13)
                zb = rng.standard_normal(m.npanels) + 1j * rng.standard_normal(m.npanels)
                zs = (rng.standard_normal(m.npanels) + 1j * rng.standard_normal(m.npanels)) * m.isym
                xeff, yeff = 3.0, -2.0
                cmd = f"mesh\n{m.npoints} {m.npanels} {m.isym} {depth:.17e}\n"
                cmd += "".join(f"{x:.17e} {y:.17e} {z:.17e}\n" for x, y, z in m.vertices)
                cmd += "".join(" ".join(str(v + 1) for v in p) + "\n" for p in m.panels)
                cmd += f"kochin\n{kw:.17e} {xeff:.17e} {yeff:.17e} {thetas.size}\n"
                cmd += "".join(f"{t:.17e}\n" for t in thetas)
                cmd += "".join(f"{a.real:.17e} {a.imag:.17e} {b.real:.17e} {b.imag:.17e}\n" for a, b in zip(zb, zs))
                lines = harness(cmd + "end\n")
                h = np.array([[float(v) for v in ln.split()] for ln in lines[-thetas.size:]])
                cases.append(dict(name=name, vertices=m.vertices, panels=m.panels, isym=m.isym, depth=depth,
                                  kwave=kw, xeff=xeff, yeff=yeff, thetas=thetas, zigb=zb, zigs=zs,
                                  h=h[:, 1] + 1j * h[:, 2]))
    out = {}
    for k, c in enumerate(cases):
        for key, v in c.items():
            out[f"c{k}_{key}"] = v
    out["ncases"] = len(cases)
    np.savez_compressed(OUT / "fn_kochin.npz", **out)
    print(f"kochin: {len(cases)} cases")


def potential_fixture() -> None:
    """COMPUTE_POTENTIAL_DOMAIN (COMPUTE_POTENTIAL_DOMAIN.f90:24-82 with
    VVV/VNV/VNVF/VVG of COMPUTE_GREEN_FREESURFACE.f90) of the reference
    build on seeded random sources: a half cylinder (y-symmetry) and a
    barge, infinite depth and h = 50 m (finite depth and one k h >= 20
    frequency), at free-surface points (z = 0) far and near the hull and at
    submerged points, some within a few panel sizes of a panel."""
    rng = np.random.default_rng(31415)
    cases = []
    cyl = nmesh.cylinder(10.0, 5.0, 16, 4, 4, half=True)
    barge = nmesh.box(100.0, 20.0, 10.0, 18, 4, 2)
    pts = {"cyl": [(11.0, 2.0, 0.0), (14.0, 9.0, 0.0), (40.0, 25.0, 0.0), (-12.0, 3.0, 0.0), (0.0, 10.8, -2.0),
                   (25.0, -30.0, -4.0), (3.0, 3.0, -6.5), (10.6, 0.5, -4.9)],
           "barge": [(51.0, 0.0, 0.0), (0.0, 10.7, 0.0), (80.0, 40.0, 0.0), (-55.0, -12.0, 0.0), (30.0, 10.9, -3.0),
                     (-120.0, 60.0, -20.0), (0.0, 0.0, -10.8), (49.0, -9.0, -10.5)]}
    for name, m in (("cyl", cyl), ("barge", barge)):
        for depth, periods in ((0.0, (4.0, 12.0)), (50.0, (2.8, 10.0, 30.0))):
            for period in periods:
                zb = rng.standard_normal(m.npanels) + 1j * rng.standard_normal(m.npanels)
                zs = (rng.standard_normal(m.npanels) + 1j * rng.standard_normal(m.npanels)) * m.isym
                p3 = np.array(pts[name], dtype=np.float64)
                cmd = f"mesh\n{m.npoints} {m.npanels} {m.isym} {depth:.17e}\n"
                cmd += "".join(f"{x:.17e} {y:.17e} {z:.17e}\n" for x, y, z in m.vertices)
                cmd += "".join(" ".join(str(v + 1) for v in p) + "\n" for p in m.panels)
                cmd += f"potential\n{period:.17e} {len(p3)}\n" + "".join(f"{a:.17e} {b:.17e} {c:.17e}\n" for a, b, c in p3)
                cmd += "".join(f"{a.real:.17e} {a.imag:.17e} {b.real:.17e} {b.imag:.17e}\n" for a, b in zip(zb, zs))
                lines = harness(cmd + "end\n")
                head = lines[-len(p3) - 1].split()
                v = np.array([[float(t) for t in ln.split()] for ln in lines[-len(p3):]])
                cases.append(dict(name=name, vertices=m.vertices, panels=m.panels, isym=m.isym, depth=depth,
                                  period=period, regime=head[0], kwave=float(head[1]), amh=float(head[2]),
                                  nexp=int(head[3]), points=p3, zigb=zb, zigs=zs, phi=v[:, 3] + 1j * v[:, 4]))
    out = {}
    for k, c in enumerate(cases):
        for key, val in c.items():
            out[f"c{k}_{key}"] = val
    out["ncases"] = len(cases)
    np.savez_compressed(OUT / "fn_potential.npz", **out)
    print(f"potential: {len(cases)} cases, regimes {[c['regime'] for c in cases]}")


FILE_CASES = {
    # name: (input files, Nemoh.cal line edits, free-surface / pressure problems kept)
    "Cylinder": (("Cylinder.dat",), {32: "0\t0.\t180.\t! Kochin", 33: "0\t10\t100.\t100.\t! FS"}, ()),
    "NonSymmetrical": (("NonSymmetrical.dat",), {32: "0\t0.\t180.\t! Kochin"}, ()),
    "2Bodies": (("Cube.dat", "Rotated_cube.dat"), {45: "4\t0.1\t2.0\t! Number of wave frequencies",
                                                   50: "0\t0.\t180.\t! Kochin"}, ()),
    # unmodified: Kochin function (181 angles), free surface (50 x 50) and
    # pressure output switched on by the reference's own input
    "Postprocessing": (("Rotated_cube.dat",), {}, (1, 7)),
}


def file_fixtures(ref: Path, tmp: Path) -> None:
    """The reference pipeline's files, as text, for the file-format layer:
    the inputs (Nemoh.cal, mesh files) and every file the preProcessor, the
    Solver and the postProcessor write (the free-surface and pressure files of the problems
    listed only, to keep the fixture small).  Stored as files_<name>.npz,
    one entry per file (its path with '/' as '|')."""
    for name, (inputs, edits, keep) in FILE_CASES.items():
        src = ref / "Verification" / name
        work = tmp / ("files_" + name)
        work.mkdir()
        for f in inputs + ("input.txt",):
            shutil.copy(src / f, work / f)
        cal = (src / "Nemoh.cal").read_text(errors="replace")
        (work / "Nemoh.cal").write_text(edit_cal(cal, edits) if edits else cal)
        run_case(work, name)
        r = subprocess.run([str(REF_BIN / "postproc")], cwd=work, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: postproc failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        out = {}
        for f in ("Nemoh.cal",) + inputs:
            out["in|" + f] = np.frombuffer((work / f).read_bytes(), dtype=np.uint8)
        files = sorted(p for p in work.rglob("*.dat") if p.parent != work or p.name == "Normalvelocities.dat")
        files += sorted(work.rglob("*.tec"))
        for p in files:
            rel = p.relative_to(work).as_posix()
            m = re.match(r"results/(freesurface|pressure)\.\s*(\d+)\.dat", rel)
            if m and int(m.group(2)) not in keep:
                continue
            out["out|" + rel.replace("/", "|")] = np.frombuffer(p.read_bytes(), dtype=np.uint8)
        np.savez_compressed(OUT / f"files_{name}.npz", **out)
        size = (OUT / f"files_{name}.npz").stat().st_size
        print(f"files {name}: {len(out)} files, {size / 1024:.0f} KiB")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    build_ref(a.ref)
    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        if a.only in ("", "fn"):
            fn_fixtures()
        if a.only in ("", "fn", "wavenumber"):
            wavenumber_fixture()
        if a.only in ("", "fn", "kochin"):
            kochin_fixture()
        if a.only in ("", "fn", "potential"):
            potential_fixture()
        if a.only in ("", "files"):
            file_fixtures(Path(a.ref), tmp)
        if a.only in ("", "pairs"):
            pair_fixtures()
        if a.only in ("", "cases"):
            verification_cases(Path(a.ref), tmp)
            synthetic_cases(tmp)
        if a.only in ("", "c4"):
            c4_bodies_cases(tmp)
        if a.only in ("", "barge4000"):
            barge4000_case(tmp)


if __name__ == "__main__":
    main()
