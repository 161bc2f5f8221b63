"""GPU parity of the on-device processing: lbm_track_errors and
lbm_velocity_stats (through the C ABI) against the oracle's sequential sums
on the same populations.

Every per-node value is formed in the oracle's order, so only the order of
the FP64 sums differs (a fixed tree on the GPU, a left fold on the CPU): the
bar is 1e-12 relative for the norms and sums, with an absolute floor scaled
by the sum of magnitudes where a sum cancels (momentum, mean velocity).
For the reduced storages the oracle reads the populations as downloaded,
which is the same FP64 reconstruction the kernel makes.
"""
import numpy as np
import pytest

from lbm_cases import QIDS, random_populations

pytestmark = pytest.mark.gpu

STORAGES = ["f64", "f32", "f16dev", "bf16dev", "f16neq", "bf16neq"]


def _L():
    import lbmhip
    return lbmhip


def _q(qid):
    return _L().Quadratures[f"D2Q{qid}"]()


def _inputs(qid, nx, ny, seed):
    f = random_populations(qid, nx, ny, seed=seed)
    rng = np.random.default_rng(seed + 100)
    E = np.asfortranarray(rng.standard_normal((nx, ny, 8)))
    aff = rng.standard_normal(16)
    return f, E, aff


def _check_track(got, ref, what):
    np.testing.assert_allclose(got[:7], ref[:7], rtol=1e-12, err_msg=what)
    np.testing.assert_allclose(got[7], ref[7], rtol=1e-12, err_msg=what)
    scale = max(abs(ref[9]), abs(ref[7])) * 1e-12
    np.testing.assert_allclose(got[8], ref[8], rtol=1e-12, atol=scale, err_msg=what)
    np.testing.assert_allclose(got[9], ref[9], rtol=1e-12, err_msg=what)
    assert got[10] == ref[10], what


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("storage", STORAGES)
def test_track_errors_matches_oracle(oracle_lib, qid, storage):
    L = _L()
    nx, ny = 37, 23
    f, E, aff = _inputs(qid, nx, ny, seed=qid)
    q = _q(qid)
    lat = L.HipLattice.from_numpy(q, f, storage=storage)
    fd = lat.to_numpy()                     # the stored state, in FP64
    ex = L.ExpectedFields(nx, ny, E)
    got = L.track_errors(lat, ex, 0.6, 0.03, 0.25, aff)
    ref = oracle_lib.track_errors(qid, fd, E, aff, 0.6, 0.03, 0.25)
    _check_track(got, ref, f"D2Q{qid} {storage}")


@pytest.mark.parametrize("shape", [(1, 1), (1, 300), (300, 1), (257, 129), (1024, 1100)])
def test_track_errors_sizes(oracle_lib, shape):
    """ragged sizes, single rows and columns, and a lattice where every thread
    of the fixed 2048-block grid visits several nodes"""
    L = _L()
    nx, ny = shape
    f, E, aff = _inputs(9, nx, ny, seed=nx + ny)
    q = _q(9)
    lat = L.HipLattice.from_numpy(q, f)
    ex = L.ExpectedFields(nx, ny, E)
    got = L.track_errors(lat, ex, 0.8, 0.01, 1.0, aff)
    ref = oracle_lib.track_errors(9, f, E, aff, 0.8, 0.01, 1.0)
    _check_track(got, ref, f"{shape}")
    # the same call twice gives the same bits
    assert np.array_equal(got, L.track_errors(lat, ex, 0.8, 0.01, 1.0, aff))


@pytest.mark.parametrize("qid", [9, 37])
@pytest.mark.parametrize("storage", ["f64", "f16neq"])
def test_velocity_stats_and_tracker(oracle_lib, qid, storage):
    L = _L()
    nx, ny = 45, 31
    q = _q(qid)
    f1 = random_populations(qid, nx, ny, seed=1)
    f2 = random_populations(qid, nx, ny, seed=2)
    a = L.HipLattice.from_numpy(q, f1, storage=storage)
    b = L.HipLattice.from_numpy(q, f2, storage=storage)
    d1, d2 = a.to_numpy(), b.to_numpy()
    tr = L.VelocityTracker(nx, ny)
    u_old = np.zeros((nx, ny, 2), order="F")
    for lat, fd in ((a, d1), (b, d2), (b, d2)):
        got = L.velocity_stats(lat, tr)
        ref = oracle_lib.velocity_stats(qid, fd, u_old)
        np.testing.assert_allclose(got[2:4], ref[2:4], rtol=1e-12, atol=1e-30)
        np.testing.assert_allclose(got[:2], ref[:2], rtol=1e-11, atol=1e-15)
        assert got[4] == ref[4] == 0
    # without a tracker only the means
    got = L.velocity_stats(a)
    ref = oracle_lib.velocity_stats(qid, d1)
    np.testing.assert_allclose(got[:2], ref[:2], rtol=1e-11, atol=1e-15)
    assert got[2] == 0 and got[3] == 0


def test_nonfinite_population_propagates():
    L = _L()
    nx, ny = 20, 9
    q = _q(9)
    f = random_populations(9, nx, ny, seed=3)
    f[4, 5, 2] = np.nan
    lat = L.HipLattice.from_numpy(q, f)
    ex = L.ExpectedFields(nx, ny, np.ones((nx, ny, 8)))
    out = L.track_errors(lat, ex, 0.7, 0.02)
    assert out[10] == 1 and np.all(np.isnan(out[:10]))
    s = L.velocity_stats(lat)
    assert s[4] == 1 and np.isnan(s[0])


def test_processing_errors():
    L = _L()
    q = _q(9)
    lat = L.HipLattice(q, 8, 6)
    with pytest.raises(L.LbmError):
        L.track_errors(lat, L.ExpectedFields(8, 7), 0.7, 0.02)
    with pytest.raises(L.LbmError):
        L.velocity_stats(lat, L.VelocityTracker(6, 8))
    with pytest.raises(L.LbmError):
        L.ExpectedFields(0, 5)
    with pytest.raises(ValueError):
        L.ExpectedFields(8, 6, np.zeros((8, 6, 7)))


def test_track_hydrodynamic_errors_over_a_decaying_vortex(oracle_lib):
    """simulate(model, 0:250) with TrackHydrodynamicErrors on the GPU records
    the row of step 250 (track_hydrodynamic_errors.jl:62-66: without
    should_process only the last step and a stop record one); it equals the
    oracle's row on the oracle's own FP64 run (bit-exact populations, so only
    the summation order differs), and the populations equal a run without
    processing.  With should_process every step records a row, the last one
    twice (the loop's next! and the one after it, lattice_boltzmann_model.jl:
    69-74)."""
    L = _L()
    pb = L.TaylorGreenVortex(scale=2, static=False)      # 32 x 32, ν = 1/6
    q = L.D2Q9()
    f0 = pb.initial_populations(q)
    tau = q.speed_of_sound_squared * pb.ν + 0.5
    pm = L.TrackHydrodynamicErrors(pb, should_process=False, n_steps=250)
    m = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(tau), processing_method=pm,
                                                 delta_t=pb.delta_t())
    L.simulate(m, range(0, 251))
    assert [r.timestep for r in pm.df] == [250]
    m2 = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(tau))
    L.simulate(m2, range(0, 251))
    assert np.array_equal(m.f_stream.to_numpy(), m2.f_stream.to_numpy())

    pm3 = L.TrackHydrodynamicErrors(pb, should_process=True, n_steps=10 ** 6)
    m3 = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(tau), processing_method=pm3,
                                                  delta_t=pb.delta_t())
    L.simulate(m3, range(0, 5))
    assert [r.timestep for r in pm3.df] == [1, 2, 3, 4, 5, 5]
    assert pm3.df[-1] == pm3.df[-2]

    planes = np.asfortranarray(pb.expected_planes(q))
    x, y = pb.range()
    delta = ((x[-1] - x[0]) / (len(x) - 1)) * ((y[-1] - y[0]) / (len(y) - 1))
    fs = np.asfortranarray(f0.copy())
    fc = np.asfortranarray(f0.copy())
    done = 0
    for row in pm3.df[:5] + pm.df:
        oracle_lib.simulate(9, fs, fc, 0, [tau], row.timestep - done)
        done = row.timestep
        aff = pb.expected_affine(q, row.timestep * pb.delta_t())
        ref = oracle_lib.track_errors(9, fs, planes, aff, q.speed_of_sound_squared * pb.ν,
                                      pb.u_max, delta)
        _check_track(np.array(list(row[1:]) + [0.0]), ref, f"t={row.timestep}")
    # the lattice solution tracks the analytic vortex: small velocity error
    assert all(r.error_u < 2e-2 for r in pm.df + pm3.df)


def test_velocity_convergence_stop_matches_oracle(oracle_lib):
    """TrackHydrodynamicErrors with a VelocityConvergenceStoppingCriteria on the
    forced (static) vortex: the device tracker's sums give the same stop step
    as the oracle's sequential sums of the same FP64 run (checks every 100
    steps, stopping_criteria.jl:71-115); the row is recorded at that step."""
    L = _L()
    pb = L.TaylorGreenVortex(scale=1, static=True)       # 16 x 16, forced
    q = L.D2Q9()
    f0 = pb.initial_populations(q)
    tau = q.speed_of_sound_squared * pb.ν + 0.5
    F = np.asfortranarray(pb.lattice_force())
    tol = 1e-6
    crit = L.VelocityConvergenceStoppingCriteria(tol, nx=pb.NX, ny=pb.NY)
    pm = L.TrackHydrodynamicErrors(pb, should_process=False, n_steps=10 ** 6, stop_criteria=crit)
    m = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(tau, F), processing_method=pm,
                                                 delta_t=pb.delta_t())
    L.simulate(m, range(0, 3000))
    assert len(pm.df) == 1
    t_stop = pm.df[0].timestep
    # the oracle's run, checked at the same steps
    S = np.asfortranarray(f0.copy())
    C = np.asfortranarray(f0.copy())
    u_old = np.zeros((pb.NX, pb.NY, 2), order="F")
    t, expect = 0, None
    while t < 3000:
        oracle_lib.simulate(9, S, C, 0, [tau], 100, force=F)
        t += 100
        s = oracle_lib.velocity_stats(9, S, u_old)
        conv = np.sqrt(s[2]) / s[3] if s[3] > 0 else np.inf
        if conv < tol:
            expect = t
            break
    assert expect == 400 and t_stop == expect
    assert np.array_equal(m.f_stream.to_numpy(), S)


@pytest.mark.parametrize("every", [3, [2, 7, 10]])
def test_take_snapshots(oracle_lib, every):
    """TakeSnapshots keeps device copies at the reference's steps
    (take_snapshots.jl:11-29), bit-identical to the oracle's states there."""
    L = _L()
    q = L.D2Q9()
    nx, ny = 20, 14
    f0 = random_populations(9, nx, ny, seed=9)
    pm = L.TakeSnapshots(None, every)
    m = L.LatticeBoltzmannModel.from_populations(q, f0, L.TRT.fields(0.7, 1.2), processing_method=pm)
    L.simulate(m, range(0, 10))                       # next! at t = 1..10, then 10 again
    want = [t for t in range(1, 11) if (t % every == 0 if isinstance(every, int) else t in every)]
    if 10 in want:
        want.append(10)                               # the call after the loop
    assert pm.timesteps == want
    S = np.asfortranarray(f0.copy())
    C = np.asfortranarray(f0.copy())
    done = 0
    for t, snap in zip(pm.timesteps, pm.arrays()):
        oracle_lib.simulate(9, S, C, 1, [0.7, 1.2], t - done)
        done = t
        assert np.array_equal(snap, S), t


@pytest.mark.parametrize("slots", [1, 3])
def test_snapshots_async_slots_and_files(oracle_lib, tmp_path, slots):
    """More snapshots than staging slots (each slot reused after its host
    copy), written as .npy files: every array equals the oracle's state at
    its step, in memory and on disk."""
    L = _L()
    q = L.D2Q9()
    nx, ny = 64, 48
    f0 = random_populations(9, nx, ny, seed=10)
    pm = L.TakeSnapshots(None, 2, device_slots=slots, directory=str(tmp_path))
    m = L.LatticeBoltzmannModel.from_populations(q, f0, L.TRT.fields(0.7, 1.2),
                                                 processing_method=pm)
    L.simulate(m, range(0, 12))
    assert pm.timesteps == [2, 4, 6, 8, 10, 12, 12]
    snaps = pm.arrays()
    S = np.asfortranarray(f0.copy())
    C = np.asfortranarray(f0.copy())
    done = 0
    for t, snap in zip(pm.timesteps, snaps):
        oracle_lib.simulate(9, S, C, 1, [0.7, 1.2], t - done)
        done = t
        assert np.array_equal(snap, S), t
        disk = np.load(tmp_path / f"snapshot_{t:010d}.npy")
        assert disk.shape == (nx, ny, 9) and disk.flags.f_contiguous
        assert np.array_equal(disk, S), t


@pytest.mark.parametrize("storage", ["f32", "f16neq", "bf16dev"])
def test_snapshots_reduced_storage(storage):
    """Snapshots of a reduced-storage field are its decoded (nx, ny, Q) Float64
    populations, as a download gives them."""
    L = _L()
    q = L.D2Q17()
    nx, ny = 40, 24
    f0 = random_populations(17, nx, ny, seed=11, noise=1e-4)
    pm = L.TakeSnapshots(None, [3, 7])
    m = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.8), processing_method=pm,
                                                 storage=storage)
    ref = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.8), storage=storage)
    L.simulate(m, range(0, 7))
    L.simulate(ref, range(0, 3))
    a3 = ref.f_stream.to_numpy()
    L.simulate(ref, range(3, 7))
    a7 = ref.f_stream.to_numpy()
    snaps = pm.arrays()
    assert pm.timesteps == [3, 7, 7]
    assert np.array_equal(snaps[0], a3)
    assert np.array_equal(snaps[1], a7) and np.array_equal(snaps[2], a7)
