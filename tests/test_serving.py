"""The request loop (nemoh_server_*) and the serving workload.

GPU tests: a request's forces are bit for bit the same alone and in a mix
of requests of other sizes, depths and precisions running concurrently;
they equal those of a sequential session over the same frequencies; the
finite-depth barge request matches the reference's forces; a failing
request leaves the others untouched.  The handle's GM (largest horizontal
vertex distance, computed on the device) equals the exact maximum.
"""
import numpy as np
import pytest

from conftest import load_golden
from nemoh_amd import bodycond as bcm
from nemoh_amd import core
from nemoh_amd import io as nio
from nemoh_amd import mesh as nmesh
from nemoh_amd import serving
from nemoh_amd.solver import Solver


def _tables():
    t = load_golden("fn_crek.npz")
    return t["xr"], t["xz"], (t["apd1x"], t["apd1z"], t["apd2x"], t["apd2z"])


def test_two_cylinder_request_shape():
    rng = np.random.default_rng(3)
    for _ in range(5):
        r = serving.two_cylinder_request(rng, n_range=(200, 3000), max_freq=4)
        nf, ntot, nrhs = r.nvel.shape
        assert r.mesh.isym == 0 and ntot == r.npanels and 150 <= ntot <= 3600
        assert nf == r.omegas.size and 13 <= nrhs <= 16 and r.nds.shape == (12, ntot)
        # each body's modes act on its own panels only
        half = ntot // 2
        assert np.all(r.nds[:6, half:] == 0) and np.all(r.nds[6:, :half] == 0)
        assert np.array_equal(r.nvel[0, :, -12:], r.nvel[-1, :, -12:])


def _barge_request():
    d = load_golden("case_barge_fd.npz")
    m = nmesh.Mesh(d["vertices"], d["panels"], int(d["isym"]))
    om = np.unique(d["omega"])
    nf = om.size
    nvel = d["nvel"].reshape(m.npanels, nf, -1).transpose(1, 0, 2)
    ref = nio.table_to_forces(d["forces_table"], d["omega"], d["kind"])
    return serving.Request(m, float(d["depth"]), om, nvel, d["nds"], float(d["rho"]), float(d["g"])), ref


def _cyl_request(seed, omegas):
    rng = np.random.default_rng(seed)
    r = serving.two_cylinder_request(rng, n_range=(500, 900), max_freq=1)
    nh = r.nvel.shape[2] - 12  # headings first, then the 12 radiation columns
    nvel = np.repeat(r.nvel[:1], len(omegas), axis=0)
    for i, w in enumerate(omegas):
        for j in range(nh):
            nvel[i, :, j] = bcm.diffraction(r.mesh, w, np.deg2rad(180.0 * j / nh), 0.0, r.g, r.rho)[0]
    return serving.Request(r.mesh, 0.0, np.asarray(omegas, float), nvel, r.nds, r.rho, r.g)


def _flat(f):
    """(nfreq, nint, nrhs) -> (nint, nfreq * nrhs) in problem order."""
    return f.transpose(1, 0, 2).reshape(f.shape[1], -1)


def _sequential(req, precision):
    s = Solver(req.mesh, req.depth, req.rho, req.g, precision=precision, tables=_tables())
    out = []
    for i, w in enumerate(req.omegas):
        _, f = s.core.solve_frequency(2 * np.pi / w, req.nvel[i], nds=req.nds, want_pressure=False)
        out.append(f)
    s.close()
    return np.stack(out)


@pytest.mark.gpu
def test_extent_matches_exact_maximum():
    rng = np.random.default_rng(11)
    for np_, isym in ((5, 0), (255, 1), (256, 0), (257, 1), (700, 0), (1100, 1)):
        v = rng.uniform(-40, 40, (np_, 3))
        v[:, 2] = -np.abs(v[:, 2]) - 1
        nq = max(1, np_ // 4)
        pan = (np.arange(4 * nq) % np_).reshape(nq, 4).astype(np.int32)
        c = core.Core(v, pan, isym, 0.0, geometry=_dummy_geometry(nq), tables=_tables())
        x, y = v[:, 0], v[:, 1]
        i, j = np.triu_indices(np_, 1)
        dx = x[j] - x[i]
        best = (dx * dx + (y[i] - y[j]) ** 2).max()
        if isym:
            best = max(best, (dx * dx + (y[i] + y[j]) ** 2).max())
        assert c.extent() == np.sqrt(best), (np_, isym)
        c.close()


def _dummy_geometry(nq):
    g = {k: np.ones(nq) for k in ("xg", "yg", "zg", "xn", "yn", "zn", "aire", "tdis")}
    g["zer"] = -1.0
    return g


@pytest.mark.gpu
def test_server_results_identical_alone_and_in_mix():
    barge, ref = _barge_request()
    a = _cyl_request(5, [0.5, 1.1, 1.7])
    b = _cyl_request(9, [0.4, 0.9])
    bad = _cyl_request(5, [0.7, 1e-5, 1.2])  # |w| < 1e-4 at its second frequency
    with serving.Server(workers=3, tables=_tables()) as srv:
        alone_a, ta = srv.wait(srv.submit(a, core.MIXED))
        alone_barge, _ = srv.wait(srv.submit(barge, core.FP64))
        print(f"alone: cylinders {a.npanels} panels, {ta['service']:.1f} ms")
        ids = [srv.submit(b, core.MIXED), srv.submit(bad, core.MIXED), srv.submit(barge, core.FP64),
               srv.submit(a, core.MIXED), srv.submit(barge, core.MIXED), srv.submit(a, core.FP64)]
        mix_b, _ = srv.wait(ids[0])
        with pytest.raises(core.NemohError) as e:
            srv.wait(ids[1])
        assert e.value.code == core.ERR_OMEGA and "frequency 1" in str(e.value)
        mix_barge, _ = srv.wait(ids[2])
        mix_a, tm = srv.wait(ids[3])
        mix_barge_mixed, _ = srv.wait(ids[4])
        mix_a_fp64, _ = srv.wait(ids[5])
    assert np.array_equal(alone_a, mix_a)
    assert np.array_equal(alone_barge, mix_barge)
    # the same as a sequential session over the request's frequencies
    assert np.array_equal(mix_a, _sequential(a, core.MIXED))
    assert np.array_equal(mix_a_fp64, _sequential(a, core.FP64))
    assert np.array_equal(mix_b, _sequential(b, core.MIXED))
    # finite-depth barge (its last frequency on the infinite-depth path with
    # the carried NQ) against the reference's forces
    def normwise(x, r):
        return float((np.abs(x - r).max(axis=0) / np.abs(r).max(axis=0)).max())
    assert normwise(_flat(mix_barge), ref) < 1e-10
    assert normwise(_flat(mix_barge_mixed), ref) < 1e-6
    assert normwise(_flat(mix_a), _flat(mix_a_fp64)) < 1e-6


def test_random_body_request_shape():
    rng = np.random.default_rng(7)
    for _ in range(12):
        r = serving.random_body_request(rng, n_range=(300, 1200), nfreq=3)
        nf, ntot, nrhs = r.nvel.shape
        assert (nf, nrhs) == (3, 7) and r.nds.shape == (6, ntot) and r.mesh.isym == 0
        assert 250 <= ntot <= 1400
        assert r.mesh.vertices[:, 2].max() <= 1e-9  # wetted surface only


@pytest.mark.gpu
def test_batched_bodies_mixed_vs_fp64():
    """A seeded batch of cylinders, boxes and spheroids through the request
    loop: mixed within 1e-6 of FP64 (normwise per problem) on every body."""
    rng = np.random.default_rng(21)
    reqs = [serving.random_body_request(rng, n_range=(300, 1500), nfreq=2) for _ in range(9)]
    out = {}
    for prec in (core.MIXED, core.FP64):
        with serving.Server(workers=4, tables=_tables()) as srv:
            ids = [srv.submit(r, prec) for r in reqs]
            out[prec] = [srv.wait(i)[0] for i in ids]
    for r, a, b in zip(reqs, out[core.MIXED], out[core.FP64]):
        err = float((np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)).max())
        assert err < 1e-6, (r.npanels, err)


@pytest.mark.gpu
def test_wave_requests_match_array_requests():
    """A request whose problems the device generates equals the same request
    with host-generated arrays (forces within 1e-12), its Froude-Krylov
    forces equal -NDS . p_incident of the host restatement, and it is the
    same alone and in a mix."""
    rng = np.random.default_rng(17)
    m, body, modes, omegas, betas = serving.two_cylinder_parts(rng, n_range=(500, 900), max_freq=3)
    arr = serving._arrays(m, body, modes, omegas, betas, 0.0, 9.81, 1025.0)
    wav = serving.WaveRequest(m, 0.0, omegas, betas, modes, None, body)
    other = serving.random_body_waves(np.random.default_rng(4), n_range=(300, 700), nfreq=2)
    with serving.Server(workers=3, tables=_tables()) as srv:
        fa, _ = srv.wait(srv.submit(arr, core.FP64))
        fw, fk, _ = srv.wait(srv.submit(wav, core.FP64), with_fk=True)
        ids = [srv.submit(other, core.MIXED), srv.submit(wav, core.FP64), srv.submit(arr, core.MIXED)]
        srv.wait(ids[0])
        fw_mix, fk_mix, _ = srv.wait(ids[1], with_fk=True)
        srv.wait(ids[2])
    # scale per frequency over all forces: the yaw of each cylinder about its
    # own axis moves no water, so those columns are rounding noise on both
    # paths and a per-column ratio would compare noise with noise
    err = max(float(np.abs(fw[i] - fa[i]).max() / np.abs(fa[i]).max()) for i in range(omegas.size))
    print(f"wave request vs array request: {err:.2e}")
    assert err < 1e-12
    assert np.array_equal(fw, fw_mix) and np.array_equal(fk, fk_mix)
    nds = bcm.integration(m, modes, body)
    for i, w in enumerate(omegas):
        pinc = np.stack([bcm.diffraction(m, w, b, 0.0, 9.81, 1025.0)[1] for b in betas], axis=1)
        ref = -(nds @ pinc)
        assert np.abs(fk[i] - ref).max() <= 1e-12 * np.abs(ref).max()



@pytest.mark.gpu
def test_budget_bounds_device_memory():
    """A budget that holds about two of the mix's largest frequencies at once
    (40 N^2 bytes each in mixed precision): every request completes with the
    forces of an unconstrained run, bit for bit; the total charged while
    several frequencies ran never exceeded the budget; and the idle handles
    hold no more than it at the end."""
    rng = np.random.default_rng(11)
    reqs = [serving.two_cylinder_request(rng, n_range=(1200, 2600), max_freq=3) for _ in range(6)]
    nmax = max(r.npanels for r in reqs)
    budget = 2.2 * 40.0 * nmax ** 2
    tables = _tables()
    out = {}
    for b in (64.0 * 2 ** 30, budget):
        with serving.Server(workers=4, budget_bytes=b, tables=tables) as srv:
            ids = [srv.submit(r, core.MIXED) for r in reqs]
            out[b] = [srv.wait(i)[0] for i in ids]
            held, peak = srv.memory()
        print(f"budget {b / 2**30:.2f} GiB: peak charged {peak / 2**30:.2f} GiB, held at the end {held / 2**30:.2f} GiB")
        if b == budget:
            assert peak <= budget and held <= budget
    for a, c in zip(out[budget], out[64.0 * 2 ** 30]):
        assert np.array_equal(a, c)


@pytest.mark.gpu
def test_batched_items_equal_single_frequencies():
    """Workers that take several frequencies of a wave request at once and
    solve them as one batch of handles (nemoh_server_set_batch) give the
    forces and Froude-Krylov forces of one frequency at a time, bit for bit:
    batches of up to 4, and of up to 8 under a budget that only holds a few
    of the mix's frequencies at once (so batches shrink, or run alone)."""
    rng = np.random.default_rng(31)
    reqs = [serving.random_body_waves(rng, n_range=(300, 1500), nfreq=5) for _ in range(4)]
    m, body, modes, omegas, betas = serving.two_cylinder_parts(rng, n_range=(800, 1600), max_freq=6)
    reqs.append(serving.WaveRequest(m, 0.0, omegas, betas, modes, None, body))
    biggest = max(r.npanels for r in reqs)
    tables = _tables()
    runs = []
    for batch, budget in ((1, None), (4, None), (8, 3 * 40.0 * biggest ** 2)):
        with serving.Server(workers=3, tables=tables, batch=batch, budget_bytes=budget) as srv:
            ids = [srv.submit(r, core.MIXED) for r in reqs]
            runs.append([srv.wait(i, with_fk=True)[:2] for i in ids])
    for k, (f1, k1) in enumerate(runs[0]):
        for other in runs[1:]:
            f2, k2 = other[k]
            assert np.all(np.isfinite(f1)) and np.array_equal(f1, f2) and np.array_equal(k1, k2), k


def test_server_batch_size_is_checked():
    """The batch-size setter rejects a null server (no device needed)."""
    assert core.load().nemoh_server_set_batch(None, 4) == core.ERR_ARG
