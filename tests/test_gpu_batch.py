"""Batches: many bodies solved together, each heavy step one launch over all
of them (nemoh_batch_*).  Every body's forces and Froude-Krylov forces must
equal, bit for bit, those of nemoh_core_solve_waves on a handle of its own:
the grouped launches give each matrix the arithmetic of its own launch
sequence.  The single-handle path itself is pinned to the reference Solver
and the oracle (test_gpu_core.py, test_gpu_c4.py).

Cases: bodies of the batched workload's generator with uneven panel counts
(odd and even numbers of 64-row blocks, one body alone in the last block
pairs) in infinite depth; boxes in finite depth (the finite-depth assembly
and the exponential images); half meshes with y-symmetry (two symmetry
blocks per body); both precisions; a frequency per body (one mesh at many
frequencies, in both depth regimes at once); a body large enough to take
the single-body assembly kernel beside grouped ones; update timing; tables
rebuilt when the degrees of freedom change; a call that fails part-way, after
which no handle claims factors its matrices no longer hold; and the batch's
fallback for a chained solve that gives up waiting (spin limit 0), which
redoes a body on its own handle.
"""
import numpy as np
import pytest

from nemoh_amd import bodycond as bcm
from nemoh_amd import core
from nemoh_amd import mesh as nmesh
from nemoh_amd import serving

pytestmark = pytest.mark.gpu


def _handle(m, depth, tables, prec, spin=None):
    c = core.Core(m.vertices, m.panels, m.isym, depth, 1025.0, 9.81, tables=tables)
    c.set_dofs(bcm.six_dof())
    c.set_precision(prec)
    if spin is not None:
        c.set_option(core.OPT_SOLVE_SPIN_LIMIT, spin)
    return c


def _check(meshes, depth, omegas, betas, prec, spin=None):
    tables = core.build_tables()
    alone = []
    for m in meshes:
        c = _handle(m, depth, tables, prec, spin)
        alone.append([c.solve_waves(w, betas) for w in omegas])
        c.close()
    cores = [_handle(m, depth, tables, prec, spin) for m in meshes]
    b = core.Batch(cores, prec, max_rhs=len(betas) + 6)
    for fi, w in enumerate(omegas):
        got = b.solve_waves(w, betas)
        for k, ((f, fk), (fa, fka)) in enumerate(zip(got, [a[fi] for a in alone])):
            assert np.all(np.isfinite(f)) and f.shape == fa.shape
            assert np.array_equal(f, fa), (k, meshes[k].npanels, w, float(np.abs(f - fa).max() / np.abs(fa).max()))
            assert np.array_equal(fk, fka), (k, w)
    b.close()
    for c in cores:
        c.close()


def _check_multi(meshes, depth, omega_rows, betas, prec):
    """Body i at omega_rows[r][i] in call r; each lone handle makes the same
    calls, so the module state a handle carries across calls (NQ) matches."""
    tables = core.build_tables()
    alone = []
    for i, m in enumerate(meshes):
        c = _handle(m, depth, tables, prec)
        alone.append([c.solve_waves(row[i], betas) for row in omega_rows])
        c.close()
    cores = [_handle(m, depth, tables, prec) for m in meshes]
    b = core.Batch(cores, prec, max_rhs=len(betas) + 6)
    for r, row in enumerate(omega_rows):
        got = b.solve_waves(np.asarray(row), betas)
        for k, (f, fk) in enumerate(got):
            fa, fka = alone[k][r]
            assert np.all(np.isfinite(f)) and f.shape == fa.shape
            assert np.array_equal(f, fa), (k, row[k], float(np.abs(f - fa).max() / np.abs(fa).max()))
            assert np.array_equal(fk, fka), (k, row[k])
    b.close()
    for c in cores:
        c.close()


def _c4_meshes(n, nrange, seed):
    rng = np.random.default_rng(seed)
    return [serving.random_body_parts(rng, nrange)[0] for _ in range(n)]


@pytest.mark.parametrize("prec", [core.MIXED, core.FP64])
def test_batch_equals_single_bodies_infinite_depth(prec):
    ms = _c4_meshes(10, (150, 1400), 11)
    blocks = sorted({(m.npanels + 63) // 64 for m in ms})
    print("panels", [m.npanels for m in ms], "blocks", blocks, flush=True)
    assert any(b % 2 for b in blocks) and any(b % 2 == 0 for b in blocks)
    _check(ms, 0.0, [0.4, 1.3], np.zeros(1), prec)


@pytest.mark.parametrize("prec", [core.MIXED, core.FP64])
def test_batch_equals_single_bodies_finite_depth(prec):
    ms = [nmesh.box(30.0, 10.0, 5.0, 15, 5, 3), nmesh.box(20.0, 8.0, 4.0, 20, 8, 4), nmesh.box(12.0, 12.0, 6.0, 6, 6, 3)]
    _check(ms, 30.0, [0.5, 1.1], np.deg2rad([0.0, 45.0]), prec)


@pytest.mark.parametrize("prec", [core.MIXED, core.FP64])
def test_batch_equals_single_bodies_symmetric(prec):
    ms = [nmesh.cylinder(5.0, 4.0, 16, 4, 3, half=True), nmesh.cylinder(8.0, 3.0, 24, 6, 5, half=True)]
    assert all(m.isym == 1 for m in ms)
    _check(ms, 0.0, [0.6, 1.5], np.zeros(1), prec)


@pytest.mark.parametrize("prec", [core.MIXED, core.FP64])
def test_batch_frequency_per_body_one_mesh(prec):
    """The frequencies of one mesh, one handle each (nemoh_batch_solve_waves_multi):
    at depth 30 m, 3.0 rad/s is past k h = 20 and takes the infinite-depth
    assembly, so the first call mixes both regimes and the second (rolled) too."""
    m = nmesh.box(30.0, 10.0, 5.0, 15, 5, 3)
    row = [0.5, 3.0, 1.1, 0.8, 2.0]
    _check_multi([m] * len(row), 30.0, [row, row[1:] + row[:1]], np.deg2rad([0.0, 30.0]), prec)


def test_batch_frequency_per_body_mixed_bodies():
    ms = _c4_meshes(4, (150, 900), 21)
    _check_multi(ms, 0.0, [[0.4, 0.9, 1.6, 2.5], [1.2, 1.2, 0.3, 0.7]], np.zeros(1), core.MIXED)


@pytest.mark.parametrize("prec", [core.MIXED, core.FP64])
def test_batch_update_timing(prec):
    """Events around the grouped rank-128 updates: one per block pair of the
    largest body per call, algorithmic flops and bytes summed over the
    bodies' launches, forces unchanged by the timing."""
    ms = [nmesh.box(30.0, 10.0, 5.0, 15, 5, 3), nmesh.box(20.0, 8.0, 4.0, 20, 8, 4)]
    tables = core.build_tables()
    cores = [_handle(m, 30.0, tables, prec) for m in ms]
    b = core.Batch(cores, prec, max_rhs=7)
    ref = b.solve_waves(0.7, np.zeros(1))
    b.set_timing(True)
    for _ in range(2):
        got = b.solve_waves(0.7, np.zeros(1))
    t_ms, flops, nbytes, launches = b.update_time()
    b.close()
    for c in cores:
        c.close()
    for (f, fk), (fr, fkr) in zip(got, ref):
        assert np.array_equal(f, fr) and np.array_equal(fk, fkr)
    nblk = [(m.npanels + 63) // 64 for m in ms]
    assert launches == 2 * len(range(0, max(nblk) - 2, 2)), (launches, nblk)
    want_f = want_b = 0.0
    for n in (m.npanels for m in ms):
        for k in range(0, (n + 63) // 64 - 2, 2):
            t = n - 64 * (k + 2)
            want_f += 8.0 * 128 * t * t
            want_b += (8 if prec == core.MIXED else 16) * (2.0 * t * t + 256.0 * t)
    assert flops == pytest.approx(2 * want_f) and nbytes == pytest.approx(2 * want_b)
    assert t_ms > 0.0


@pytest.mark.parametrize("prec", [core.MIXED, core.FP64])
def test_batch_large_body_own_assembly(prec):
    """A body whose assembly fills the GPU on its own (1664 panels) takes the
    single-body kernel inside the batch, the small one the grouped launch:
    both equal the lone handles, at one frequency and at a frequency each in
    both depth regimes (3.0 rad/s is infinite depth at 30 m)."""
    big, small = nmesh.box(40.0, 20.0, 10.0, 32, 16, 12), nmesh.box(12.0, 12.0, 6.0, 6, 6, 3)
    assert big.npanels == 1664
    _check([big, small], 30.0, [0.8], np.deg2rad([0.0, 30.0]), prec)
    _check_multi([big, small, big], 30.0, [[0.8, 3.0, 3.0], [3.0, 0.8, 1.1]], np.zeros(1), prec)


def test_batch_rebuilds_after_dofs_change():
    """Degrees of freedom changed between two calls of one batch, with the
    same number of problems (6 dofs + 1 heading, then 3 dofs + 4 headings):
    the batch rebuilds its tables and both calls equal the lone handles."""
    ms = [nmesh.box(30.0, 10.0, 5.0, 15, 5, 3), nmesh.box(20.0, 8.0, 4.0, 20, 8, 4)]
    tables = core.build_tables()
    plans = [(bcm.six_dof(), np.zeros(1)), (bcm.six_dof()[:3], np.deg2rad([0.0, 30.0, 60.0, 90.0]))]
    alone = []
    for m in ms:
        c = _handle(m, 30.0, tables, core.MIXED)
        row = []
        for dofs, betas in plans:
            c.set_dofs(dofs)
            row.append(c.solve_waves(0.9, betas))
        alone.append(row)
        c.close()
    cores = [_handle(m, 30.0, tables, core.MIXED) for m in ms]
    b = core.Batch(cores, core.MIXED, max_rhs=7)
    for p, (dofs, betas) in enumerate(plans):
        for c in cores:
            c.set_dofs(dofs)
        got = b.solve_waves(0.9, betas)
        for k, (f, fk) in enumerate(got):
            assert np.array_equal(f, alone[k][p][0]) and np.array_equal(fk, alone[k][p][1]), (p, k)
    b.close()
    for c in cores:
        c.close()


def test_batch_failed_call_invalidates_cached_factors():
    """A call that fails part-way (here: the second body solves a different
    number of problems, found after the first body's own assembly launch was
    queued) leaves no handle claiming factors that its matrices no longer
    hold: the first body, solved alone again at its earlier frequency, gives
    its earlier forces bit for bit."""
    tables = core.build_tables()
    big = _handle(nmesh.box(40.0, 20.0, 10.0, 32, 16, 12), 30.0, tables, core.MIXED)
    small = _handle(nmesh.box(12.0, 12.0, 6.0, 6, 6, 3), 30.0, tables, core.MIXED)
    small.set_dofs(bcm.six_dof()[:3])
    betas = np.zeros(1)
    before = big.solve_waves(0.8, betas)
    b = core.Batch([big, small], core.MIXED, max_rhs=7)
    with pytest.raises(core.NemohError) as e:
        b.solve_waves(np.array([1.1, 1.1]), betas)
    assert e.value.code == core.ERR_ARG
    again = big.solve_waves(0.8, betas)
    assert np.array_equal(again[0], before[0]) and np.array_equal(again[1], before[1])
    b.close()
    big.close()
    small.close()


def test_batch_chained_solve_fallback():
    """Spin limit 0: every chained wait gives up at once, and each body is
    solved again on its own handle (stepwise), as a handle alone does."""
    ms = _c4_meshes(3, (150, 700), 5)
    _check(ms, 0.0, [0.9], np.zeros(1), core.MIXED, spin=0)


def test_batch_argument_errors():
    tables = core.build_tables()
    a = _handle(nmesh.cylinder(5.0, 4.0, 16, 4, 3, half=True), 0.0, tables, core.FP64)
    b = _handle(nmesh.box(10.0, 5.0, 3.0, 5, 3, 2), 0.0, tables, core.FP64)
    with pytest.raises(core.NemohError) as e:
        core.Batch([a, b], core.FP64)  # symmetric and unsymmetric bodies
    assert e.value.code == core.ERR_ARG
    with pytest.raises(core.NemohError):
        core.Batch([b, b], core.FP64)  # the same handle twice
    bt = core.Batch([b], core.FP64, max_rhs=7)
    with pytest.raises(core.NemohError):
        bt.solve_waves(0.8, np.zeros(3))  # 3 headings + 6 dofs > max_rhs
    f = bt.solve_waves(0.8, np.zeros(1))
    assert np.all(np.isfinite(f[0][0]))
    with pytest.raises(ValueError):
        bt.solve_waves([0.8, 0.9], np.zeros(1))  # two frequencies for one body
    with pytest.raises(core.NemohError) as e:
        bt.solve_waves([0.0], np.zeros(1))
    assert e.value.code == core.ERR_OMEGA
    bt.close()
    a.close()
    b.close()
