"""GPU parity: the gfx950 kernels (through the C ABI) against the CPU oracle.

Bar: the plain FP64 path is bit-exact against the oracle (same evaluation
order, no FMA contraction).  The required tolerance for the FP64 path is
1e-12 relative per population after N steps; asserting equality is stricter.
"""
import numpy as np
import pytest

from lbm_cases import (BC_NAMES, QIDS, bc_dicts, force_for, random_populations,
                       shear_wave_fields, taylor_green_vortex_pops, to_lbmhip)

pytestmark = pytest.mark.gpu


def _L():
    import lbmhip
    return lbmhip


def _q(qid):
    L = _L()
    return L.Quadratures[f"D2Q{qid}"]()


def _cm(kind, qid, force=None):
    L = _L()
    if kind == 0:
        return L.SRT(0.8, force), [0.8]
    if kind == 1:
        return L.TRT.fields(0.7, 1.3, force), [0.7, 1.3]
    q = _q(qid)
    N = max(2, L.order(q) // 2)
    taus = [0.9, 0.75, 1.1, 0.6][:max(N, 2)]
    return L.MRT(q, taus, force=force), taus


def _assert_same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        rel = d.max() / np.abs(b).max()
        raise AssertionError(f"{what}: not bit-exact, max |diff| {d.max():.3e} (rel {rel:.3e}), "
                             f"{np.count_nonzero(d)} of {d.size} differ")


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("force", ["none", "const", "field"])
def test_collide_bit_exact(oracle_lib, qid, coll, force):
    L = _L()
    nx, ny = 13, 7
    f = random_populations(qid, nx, ny, seed=qid + 10 * coll)
    F = force_for(force, nx, ny)
    cm, taus = _cm(coll, qid, F)
    q = _q(qid)
    fin = L.HipLattice.from_numpy(q, f)
    fout = L.HipLattice(q, nx, ny)
    L.collide_(cm, q, fin, fout, time=0.0)
    got = fout.to_numpy()
    ref = np.zeros_like(f, order="F")
    if coll == 0:
        oracle_lib.collide_srt(qid, f, ref, taus[0], F)
    elif coll == 1:
        oracle_lib.collide_trt(qid, f, ref, taus[0], taus[1], F)
    else:
        oracle_lib.collide_mrt(qid, f, ref, taus, F)
    _assert_same(got, ref, f"collide D2Q{qid} coll={coll} force={force}")
    # the input is left untouched
    _assert_same(fin.to_numpy(), f, "collide input")


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("shape", [(11, 6), (1, 1), (2, 3), (1, 5)])
def test_stream_bit_exact(oracle_lib, qid, shape):
    L = _L()
    nx, ny = shape
    f = random_populations(qid, nx, ny, seed=3)
    q = _q(qid)
    a = L.HipLattice.from_numpy(q, f)
    b = L.HipLattice(q, nx, ny)
    L.stream_(q, a, b)
    ref = np.zeros_like(f, order="F")
    oracle_lib.stream(qid, f, ref)
    _assert_same(b.to_numpy(), ref, f"stream D2Q{qid} {shape}")


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("bcname", BC_NAMES[1:])
def test_apply_bcs_bit_exact(oracle_lib, qid, bcname):
    L = _L()
    nx, ny = 9, 8
    fnew = random_populations(qid, nx, ny, seed=4)
    fold = random_populations(qid, nx, ny, seed=5)
    bcs = bc_dicts(bcname, nx, ny)
    q = _q(qid)
    dn = L.HipLattice.from_numpy(q, fnew)
    do = L.HipLattice.from_numpy(q, fold)
    L.apply_(to_lbmhip(bcs), q, dn, do)
    ref = fnew.copy(order="F")
    oracle_lib.apply_bcs(qid, bcs, ref, fold)
    _assert_same(dn.to_numpy(), ref, f"apply! D2Q{qid} {bcname}")


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("bcname", BC_NAMES)
def test_simulate_bit_exact(oracle_lib, qid, coll, bcname):
    """Fused pull kernel: both model buffers equal the reference step loop."""
    L = _L()
    nx, ny = 12, 10
    f0 = random_populations(qid, nx, ny, seed=7)
    F = force_for("const" if bcname == "poiseuille" else "none", nx, ny)
    cm, taus = _cm(coll, qid, F)
    bcs = bc_dicts(bcname, nx, ny)
    q = _q(qid)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs))
    nsteps = 17
    L.simulate(model, range(nsteps))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, nsteps, force=F, bcs=bcs)
    _assert_same(model.f_stream.to_numpy(), S, f"f_stream D2Q{qid} coll={coll} {bcname}")
    _assert_same(model.f_collision.to_numpy(), C, f"f_collision D2Q{qid} coll={coll} {bcname}")


@pytest.mark.parametrize("qid", [9, 17, 37])
@pytest.mark.parametrize("shape", [(1, 5), (3, 5), (2, 2), (1, 1), (5, 3)])
def test_tiny_lattices(oracle_lib, qid, shape):
    """|c| >= n wraps (stream.jl:69-74) and BCs hitting most rows (nx = 1 as
    CouetteFlow, couette_flow.jl:19; NX = 3 as PoiseuilleFlow, poiseuille.jl:22)."""
    L = _L()
    nx, ny = shape
    f0 = random_populations(qid, nx, ny, seed=11)
    for bcname in ["periodic", "couette", "poiseuille"]:
        bcs = bc_dicts(bcname, nx, ny)
        cm, taus = _cm(1, qid)
        q = _q(qid)
        model = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs))
        L.simulate(model, range(6))
        S = f0.copy(order="F")
        C = f0.copy(order="F")
        oracle_lib.simulate(qid, S, C, 1, taus, 6, bcs=bcs)
        _assert_same(model.f_stream.to_numpy(), S, f"tiny {shape} D2Q{qid} {bcname}")


def test_separate_operators_match_fused(oracle_lib):
    """collide_ + stream_ + apply_ (the model-level calls) == simulate."""
    L = _L()
    qid, nx, ny = 9, 16, 12
    f0 = random_populations(qid, nx, ny, seed=2)
    bcs = to_lbmhip(bc_dicts("couette", nx, ny))
    q = _q(qid)
    a = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.9), bcs)
    b = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.9), bcs)
    for t in range(5):
        L.collide_model_(a, time=float(t))
        L.stream_model_(a)
        L.apply_boundary_conditions_(a, time=float(t))
    L.simulate(b, range(5))
    _assert_same(a.f_stream.to_numpy(), b.f_stream.to_numpy(), "separate vs fused f_stream")
    _assert_same(a.f_collision.to_numpy(), b.f_collision.to_numpy(), "separate vs fused f_coll")


def test_callable_force_per_step(oracle_lib):
    """A time-dependent force closure (x_idx, y_idx, t) is re-evaluated each step."""
    L = _L()
    qid, nx, ny = 9, 8, 6
    f0 = random_populations(qid, nx, ny, seed=9)

    def force(x, y, t):
        return (1e-5 * np.cos(0.3 * y - 0.1 * t), 2e-6 * np.sin(0.2 * x))

    q = _q(qid)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.8, force), delta_t=1.0)
    L.simulate(model, range(4))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    for t in range(4):
        F = np.empty((nx, ny, 2), order="F")
        for x in range(nx):
            for y in range(ny):
                F[x, y] = force(x + 1, y + 1, float(t))
        oracle_lib.simulate(qid, S, C, 0, [0.8], 1, force=F)
    _assert_same(model.f_stream.to_numpy(), S, "time-dependent force")


def test_shear_wave_c1(oracle_lib):
    """C1: D2Q9 SRT, periodic 64x64, tau = 0.8, u_x = 0.01 sin(2 pi y/64),
    equilibrium init, 1000 steps.  GPU == oracle, and the decay of the mode
    matches exp(-nu k^2 t), nu = (tau - 1/2)/3, to 1e-3."""
    L = _L()
    n, tau, steps = 64, 0.8, 1000
    q = L.D2Q9()
    rho, ux, uy = shear_wave_fields(n)
    f0 = L.equilibrium_populations(q, rho, ux, uy)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(tau))
    L.simulate(model, range(steps))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.simulate(9, S, C, 0, [tau], steps)
    got = model.f_stream.to_numpy()
    rel = np.abs(got - S).max() / np.abs(S).max()
    assert rel <= 1e-12, rel
    _assert_same(got, S, "C1 shear wave")
    m = L.moments(model.f_stream, tau, stress=False)
    y = np.arange(n)
    amp = 2.0 / n * np.sum(m.ux[0] * np.sin(2 * np.pi * y / n))
    nu = (tau - 0.5) / 3.0
    k = 2 * np.pi / n
    expect = 0.01 * np.exp(-nu * k * k * steps)
    # the LBM shear mode decays at nu k^2 (1 + O(k^2)); k = 2 pi / 64 gives ~1e-3
    assert abs(amp / expect - 1.0) < 2e-3, (amp, expect)


def test_taylor_green_trt_256(oracle_lib):
    """C2 at N = 256 (fewer steps): D2Q9 TRT Lambda = 1/4, tau = 0.6."""
    L = _L()
    n, tau, steps = 256, 0.6, 200
    q = L.D2Q9()
    f0 = taylor_green_vortex_pops(q, n)
    cm = L.trt_magic(tau, 0.25)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm)
    L.simulate(model, range(steps))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.set_threads(8)
    oracle_lib.simulate(9, S, C, 1, [cm.τ_symmetric, cm.τ_asymmetric], steps)
    oracle_lib.set_threads(1)
    _assert_same(model.f_stream.to_numpy(), S, "TGV 256 TRT")


@pytest.mark.parametrize("qid", QIDS)
def test_moments_and_sums(oracle_lib, qid):
    L = _L()
    nx, ny = 19, 7
    f = random_populations(qid, nx, ny, seed=21)
    q = _q(qid)
    d = L.HipLattice.from_numpy(q, f)
    m = L.moments(d, 0.7)
    ref = oracle_lib.moments(qid, f, 0.7)
    for a, b, name in zip(m, ref, ["rho", "ux", "uy", "sxx", "sxy", "syy"]):
        _assert_same(a, b, f"moments {name} D2Q{qid}")
    s = L.conserved_sums(d)
    rho, ux, uy = ref[0], ref[1], ref[2]
    np.testing.assert_allclose(s.mass, rho.sum(), rtol=1e-14)
    np.testing.assert_allclose(s.momentum_x, (rho * ux).sum(), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(s.momentum_y, (rho * uy).sum(), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(s.energy, (rho * (ux * ux + uy * uy)).sum(), rtol=1e-12)
    assert s.nonfinite_nodes == 0


def test_nonfinite_detection():
    """A NaN written from the host is counted by the reduction (the NaN stop
    criterion of stopping_criteria.jl:47-50)."""
    L = _L()
    q = L.D2Q9()
    f = random_populations(9, 8, 8, seed=1)
    f[3, 4, 2] = np.nan
    f[5, 1, 0] = np.inf
    d = L.HipLattice.from_numpy(q, f)
    assert L.conserved_sums(d).nonfinite_nodes == 2


def test_abi_errors():
    L = _L()
    q = L.D2Q9()
    a = L.HipLattice(q, 8, 8)
    with pytest.raises(L.LbmError, match="distinct"):
        L.collide_(L.SRT(0.8), q, a, a)
    b = L.HipLattice(L.D2Q17(), 8, 8)
    with pytest.raises(ValueError, match="populations"):
        L.stream_(q, a, b)
    c = L.HipLattice(q, 8, 8, storage="f32")
    with pytest.raises(L.LbmError):
        L.stream_(q, a, c)
    with pytest.raises(L.LbmError, match="outside"):
        L.apply_([L.BounceBack(L.North(), (1, 9), (1, 8))], q, a, L.HipLattice(q, 8, 8))


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("nt", [0, 1])
@pytest.mark.parametrize("bcname", ["periodic", "cavity", "partial"])
def test_vectorized_fused_bit_exact(oracle_lib, qid, nt, bcname):
    """Two nodes per thread (vector loads of shifted planes, vector stores,
    optionally non-temporal) give the same bits as the scalar path."""
    L = _L()
    nx, ny = 16, 10
    f0 = random_populations(qid, nx, ny, seed=13)
    cm, taus = _cm(1, qid, force_for("const", nx, ny))
    bcs = bc_dicts(bcname, nx, ny)
    q = _q(qid)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs))
    model.launch = (64, 2, 2, nt)   # 64 x 2 threads, 2 nodes per thread
    L.simulate(model, range(9))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, 1, taus, 9, force=force_for("const", nx, ny), bcs=bcs)
    _assert_same(model.f_stream.to_numpy(), S, f"vec2 nt={nt} D2Q{qid} {bcname}")
    _assert_same(model.f_collision.to_numpy(), C, f"vec2 nt={nt} D2Q{qid} {bcname} (C)")


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("launch", [(64, 2, 1, 1), (128, 1, 1, 1), (64, 1, 1, 0)])
@pytest.mark.parametrize("bcname", ["periodic", "cavity"])
def test_mrt_launch_shapes_bit_exact(oracle_lib, qid, launch, bcname):
    """MRT through the row-shaped kernel with non-temporal stores (nt = 1) and
    through the plain one-node kernel (nt = 0), with a field force and BCs:
    the oracle's bits either way."""
    L = _L()
    nx, ny = 21, 10
    f0 = random_populations(qid, nx, ny, seed=17)
    F = force_for("field", nx, ny)
    cm, taus = _cm(2, qid, F)
    bcs = bc_dicts(bcname, nx, ny)
    q = _q(qid)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs))
    model.launch = launch
    L.simulate(model, range(6))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, 2, taus, 6, force=F, bcs=bcs)
    _assert_same(model.f_stream.to_numpy(), S, f"MRT {launch} D2Q{qid} {bcname}")
    _assert_same(model.f_collision.to_numpy(), C, f"MRT {launch} D2Q{qid} {bcname} (C)")


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("nx", [32, 33])
def test_row_shaped_and_flat_kernels_bit_exact(oracle_lib, qid, coll, nx):
    """One node per thread: rows of whole 128-byte lines (nx = 32 in FP64) run
    the row-shaped kernel, other widths (nx = 33) the flat-index kernel; with
    a field force and the cavity BC list both give the oracle's bits."""
    L = _L()
    ny = 9
    f0 = random_populations(qid, nx, ny, seed=23 + coll)
    F = force_for("field", nx, ny)
    cm, taus = _cm(coll, qid, F)
    bcs = bc_dicts("cavity", nx, ny)
    q = _q(qid)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs))
    model.launch = (64, 1, 1, 1)
    L.simulate(model, range(5))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, 5, force=F, bcs=bcs)
    _assert_same(model.f_stream.to_numpy(), S, f"nx={nx} coll={coll} D2Q{qid}")
    _assert_same(model.f_collision.to_numpy(), C, f"nx={nx} coll={coll} D2Q{qid} (C)")


def test_vec_requires_divisible_nx():
    L = _L()
    q = L.D2Q9()
    f0 = random_populations(9, 15, 6, seed=1)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.8))
    model.launch = (64, 2, 0)
    with pytest.raises(L.LbmError):
        L.simulate(model, range(3))


# ---- reduced-precision storages -------------------------------------------------
# random velocity fields are far from equilibrium (|f - feq| ~ 0.03 w), the
# worst case for the *neq storages; smooth flows are checked below
_LP_TOL = {"f32": 2e-5, "f16dev": 5e-3, "bf16dev": 5e-2, "f16neq": 1e-4, "bf16neq": 5e-4}


def _max_vec(Q, storage):
    """Largest nodes per thread the library instantiates (max_vec of
    lbm_kernels.h): rows for the non-equilibrium storages, x-runs otherwise."""
    if storage.endswith("neq"):
        return 16
    cap = 4 if storage == "f32" else 8
    return cap if Q <= 9 else (min(cap, 4) if Q <= 17 else 2)


_LP_CASES = [(qid, st, vec) for qid in (9, 17, 37)
             for st in ("f32", "f16dev", "bf16dev", "f16neq", "bf16neq")
             for vec in (1, 2, 4, 8, 16)
             if vec <= _max_vec({9: 9, 17: 17, 37: 37}[qid], st)]


@pytest.mark.parametrize("qid,storage,vec", _LP_CASES)
@pytest.mark.parametrize("coll", [0, 1])
def test_reduced_storage_tracks_oracle(oracle_lib, qid, storage, vec, coll):
    """FP32 arithmetic on FP32 / FP16-deviation / BF16-deviation /
    non-equilibrium storage stays within the storage's precision of the
    oracle's FP64 run (relative L2 velocity error after 20 steps of a Couette
    + force case), for every launch shape the library instantiates."""
    from oracle import lbm_numpy as ln
    L = _L()
    q = _q(qid)
    nx, ny = 32, 12
    f0 = random_populations(qid, nx, ny, seed=17, noise=1e-4)
    F = force_for("const", nx, ny)
    cm, taus = _cm(coll, qid, F)
    bcd = bc_dicts("couette", nx, ny)
    S, C = f0.copy(order="F"), f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, 20, force=F, bcs=bcd)
    rho, ux, uy = ln.moments_field(ln.by_id(qid), S)
    lp = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcd), storage=storage)
    lp.launch = (64, vec, 0)
    L.simulate(lp, range(20))
    b = L.moments(lp.f_stream, stress=False)
    num = np.sqrt(np.sum((ux - b.ux) ** 2 + (uy - b.uy) ** 2))
    den = np.sqrt(np.sum(ux ** 2 + uy ** 2))
    assert num / den < _LP_TOL[storage], (storage, num / den)
    np.testing.assert_allclose(b.rho, rho, rtol=_LP_TOL[storage] * 0.1)


@pytest.mark.parametrize("storage", ["f32", "f16dev", "bf16dev", "f16neq", "bf16neq"])
def test_reduced_storage_roundtrip(storage):
    L = _L()
    q = L.D2Q9()
    f = random_populations(9, 24, 8, seed=2)
    d = L.HipLattice.from_numpy(q, f, storage=storage)
    g = d.to_numpy()
    w = q.weights
    dev_err = np.abs((g - f) / w).max()
    # |f/w - 1| reaches ~0.2 here: half an ulp of that in binary16 is 1.2e-4,
    # in bfloat16 7.8e-4, in binary32 6e-9 (relative to the weight)
    # the *neq storages round only f - feq(rho, u) (~2e-3 w here) to 16 bits
    bound = {"f32": 1e-7, "f16dev": 2.5e-4, "bf16dev": 2e-3, "f16neq": 1e-5,
             "bf16neq": 5e-5}[storage]
    assert dev_err < bound, (storage, dev_err)


def test_autotune_keeps_state(oracle_lib):
    L = _L()
    q = L.D2Q9()
    n = 128
    f0 = taylor_green_vortex_pops(q, n)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, L.trt_magic(0.6))
    best = L.autotune(model, trial_steps=3)
    assert best["block"] in (64, 128, 256) and best["nodes_per_thread"] in (1, 2)
    assert best["block"] * best["rows_per_block"] <= 256
    assert best["us_per_step"] > 0
    _assert_same(model.f_stream.to_numpy(), f0, "autotune left the state untouched")
    L.simulate(model, range(5))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.simulate(9, S, C, 1, [model.collision_model.τ_symmetric,
                                      model.collision_model.τ_asymmetric], 5)
    _assert_same(model.f_stream.to_numpy(), S, "tuned launch shape")


def test_autotune_validates_model():
    """The autotuner runs the fused kernel on the model's force field and BC
    list: it applies the same checks as simulate before any launch."""
    import ctypes

    import lbmhip._lib as lib

    L = _L()
    q = L.D2Q9()
    f0 = taylor_green_vortex_pops(q, 32)
    # through the C ABI: a model whose force field belongs to a 16 x 32 lattice
    C = lib.lib()
    m = ctypes.c_void_p()
    assert C.lbm_model_create(ctypes.byref(m), 9) == 0
    try:
        assert C.lbm_model_set_srt(m, 0.8) == 0
        wrong = np.zeros((16, 32, 2), order="F")
        assert C.lbm_model_set_force(m, lib.FORCE_FIELD, lib.dptr(wrong), 16, 32) == 0
        f = L.HipLattice.from_numpy(q, f0)
        b, r, v, n = (ctypes.c_int() for _ in range(4))
        us = ctypes.c_double()
        rc = C.lbm_autotune(m, f.handle, 1, ctypes.byref(b), ctypes.byref(r), ctypes.byref(v),
                            ctypes.byref(n), ctypes.byref(us))
        assert rc != 0 and b"force field" in C.lbm_last_error()
    finally:
        C.lbm_model_destroy(m)
    bcs = [L.BounceBack(L.North(), (1, 33), (1, 32))]
    model = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.8), bcs)
    with pytest.raises(L.LbmError, match="outside"):
        L.autotune(model, trial_steps=1)
    u = np.zeros((32, 32, 2), order="F")
    model = L.LatticeBoltzmannModel.from_populations(
        q, f0, L.IterativeInitializationCollisionModel(1.0, u))
    with pytest.raises(L.LbmError, match="fused step"):
        L.autotune(model, trial_steps=1)


@pytest.mark.parametrize("storage", ["f16neq", "bf16neq"])
@pytest.mark.parametrize("qid", [9, 37])
def test_neq_storage_operators(storage, qid):
    """collide_, stream_, apply_ and the fused step agree on the
    non-equilibrium storages (each node re-encoded about its own moments)."""
    L = _L()
    q = _q(qid)
    nx, ny = 16, 10
    f0 = random_populations(qid, nx, ny, seed=23, noise=1e-4)
    bcs = to_lbmhip(bc_dicts("cavity", nx, ny))
    a = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.9), bcs, storage=storage)
    b = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.9), bcs, storage=storage)
    ref = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(0.9), bcs)
    for t in range(4):
        L.collide_model_(a, time=float(t))
        L.stream_model_(a)
        L.apply_boundary_conditions_(a, time=float(t))
    L.simulate(b, range(4))
    L.simulate(ref, range(4))
    fa, fb, fr = a.f_stream.to_numpy(), b.f_stream.to_numpy(), ref.f_stream.to_numpy()
    tol = 2e-5 if storage == "f16neq" else 1e-4
    assert np.abs(fa - fr).max() < tol and np.abs(fb - fr).max() < tol


_SMOOTH_TOL = {"f32": 1e-5, "f16neq": 5e-5, "bf16neq": 2e-4, "f16dev": 1e-2, "bf16dev": 5e-2}


@pytest.mark.parametrize("storage", sorted(_SMOOTH_TOL))
def test_reduced_storage_taylor_green(storage):
    """Smooth Taylor-Green flow, D2Q9 TRT 256^2, 500 steps: relative L2
    velocity error of each reduced storage against the FP64 run."""
    L = _L()
    q = L.D2Q9()
    f0 = taylor_green_vortex_pops(q, 256)
    cm = L.trt_magic(0.6)
    ref = L.LatticeBoltzmannModel.from_populations(q, f0, cm)
    lp = L.LatticeBoltzmannModel.from_populations(q, f0, cm, storage=storage)
    L.simulate(ref, range(500))
    L.simulate(lp, range(500))
    a = L.moments(ref.f_stream, stress=False)
    b = L.moments(lp.f_stream, stress=False)
    err = np.sqrt(np.sum((a.ux - b.ux) ** 2 + (a.uy - b.uy) ** 2) / np.sum(a.ux ** 2 + a.uy ** 2))
    assert err < _SMOOTH_TOL[storage], (storage, err)


@pytest.mark.parametrize("storage", ["f16neq", "bf16neq"])
@pytest.mark.parametrize("rows", [2, 4, 8, 16])
@pytest.mark.parametrize("shape", [(3, 3), (4, 5), (1, 6), (7, 1), (2, 9), (6, 5), (10, 3),
                                   (33, 19), (64, 21), (130, 7), (514, 9)])
@pytest.mark.parametrize("bcname", ["periodic", "couette"])
def test_neq_row_marching_matches_single_node(storage, rows, shape, bcname):
    """The row-marching kernel (interior) + edge kernel against the one node
    per thread kernel, for lattices with fewer rows than a thread marches,
    one-node-wide lattices and heights that are not multiples of the row count.
    Even widths above 4 run the two-columns-per-lane kernel: 6 is the
    narrowest (one interior pair per row), and 130 / 514 leave a single pair
    in the last block of 64 / 256 threads; odd widths run one column per
    lane.
    The two evaluate the same expressions in a different instruction order, so
    they agree bit for bit until an FP32 rounding difference flips one 16-bit
    rounding; from then on they differ by 16-bit ulps of the non-equilibrium
    parts (~1e-5 w here for binary16, ~3e-4 w for bfloat16) at a few nodes
    (measured: under 5% of the values after 7 steps).  A wrong record or edge
    mapping moves whole neighbourhoods by w s du ~ 1e-2 w."""
    L = _L()
    q = L.D2Q9()
    nx, ny = shape
    f0 = random_populations(9, nx, ny, seed=31, noise=1e-4)
    cm = L.TRT.fields(0.7, 1.3, force_for("const", nx, ny))
    bcs = to_lbmhip(bc_dicts(bcname, nx, ny))
    ulp_tol = {"f32": 1e-5, "f16dev": 1e-4, "bf16dev": 1e-3, "f16neq": 1e-4,
               "bf16neq": 1e-3}[storage]  # bfloat16: 8x coarser than binary16
    for steps, tol, frac in ((1, ulp_tol, 0.1), (7, ulp_tol, 0.1)):
        out = []
        # (block, rows per block, nodes per thread, NT): one node per thread,
        # row marching with 64- and 256-thread blocks
        for launch in ((64, 1, 1, 1), (64, 1, rows, 1), (256, 1, rows, 1)):
            m = L.LatticeBoltzmannModel.from_populations(q, f0, cm, bcs, storage=storage)
            m.launch = launch
            L.simulate(m, range(steps))
            out.append((m.f_stream.to_numpy(), m.f_collision.to_numpy()))
        for other in out[1:]:
            for a, b in zip(out[0], other):
                d = np.abs((a - b) / q.weights)
                assert d.max() <= tol, (steps, d.max())
                assert np.count_nonzero(d > 1e-6) <= frac * d.size, (steps,
                                                                     np.count_nonzero(d > 1e-6))


@pytest.mark.parametrize("qid", [13, 17, 21, 37])
@pytest.mark.parametrize("storage", ["f16neq", "bf16neq"])
@pytest.mark.parametrize("rows", [2, 4, 8, 16])
@pytest.mark.parametrize("shape", [(5, 5), (7, 9), (9, 7), (40, 23), (70, 31)])
@pytest.mark.parametrize("bcname", ["periodic", "couette"])
def test_neq_row_marching_wide_stencils(qid, storage, rows, shape, bcname):
    """The row-marching kernel on stencils reaching 2 (D2Q13) or 3 (D2Q17/21/37)
    nodes: an LDS window of 2M + 1 record rows and M halo columns, and an edge
    frame M nodes wide.  Against the one node per thread kernel, with the same
    rounding-flip allowance as for D2Q9; lattices no wider or taller than 2M
    have no interior and take the edge kernel alone.  After 7 steps the
    rounding flips have spread further through the wider neighbourhoods
    (measured: 1.2e-4 w at most, D2Q17 FP16-NEQ 70 x 31), so the allowance is
    three times the one-step bound there; a wrong window row or halo column
    moves populations by ~1e-2 w."""
    L = _L()
    q = _q(qid)
    nx, ny = shape
    f0 = random_populations(qid, nx, ny, seed=37, noise=1e-4)
    cm = L.TRT.fields(0.7, 1.3, force_for("const", nx, ny))
    bcs = to_lbmhip(bc_dicts(bcname, nx, ny))
    ulp_tol = {"f32": 1e-5, "f16dev": 1e-4, "bf16dev": 1e-3, "f16neq": 1e-4,
               "bf16neq": 1e-3}[storage]
    w = np.asarray(q.weights, dtype=float)
    w = np.where(w > 0, w, 1.0)                 # D2Q21's four zero-weight planes stay 0
    # (on the smallest lattices the 7-step flips reach up to a quarter of the
    # values: D2Q37 on 5 x 5, 13% measured)
    for steps, tol, frac in ((1, ulp_tol, 0.1), (7, 3 * ulp_tol, 0.25)):
        out = []
        for launch in ((64, 1, 1, 1), (64, 1, rows, 1), (256, 1, rows, 1)):
            m = L.LatticeBoltzmannModel.from_populations(q, f0, cm, bcs, storage=storage)
            m.launch = launch
            L.simulate(m, range(steps))
            out.append((m.f_stream.to_numpy(), m.f_collision.to_numpy()))
        for other in out[1:]:
            for a, b in zip(out[0], other):
                d = np.abs((a - b) / w)
                assert d.max() <= tol, (steps, d.max())
                assert np.count_nonzero(d > 1e-6) <= frac * d.size, (steps,
                                                                     np.count_nonzero(d > 1e-6))


@pytest.mark.parametrize("qid", [4, 5, 9])
@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("bcname,force", [("periodic", "none"), ("couette", "const"),
                                          ("poiseuille", "const"), ("cavity", "field"),
                                          ("partial", "none")])
@pytest.mark.parametrize("steps", [2, 5])
def test_first_step_tile_bit_exact(oracle_lib, qid, coll, bcname, force, steps):
    """FP64 one-speed stencils on lattices of whole 64 x 8 tiles start a run
    with one pass (collide S_0 into an LDS tile with a one-node halo, then
    stream, boundaries and collide): f_stream and f_collision equal the
    oracle's bits, with boundaries on every side and with forces."""
    L = _L()
    nx, ny = 128, 16
    f0 = random_populations(qid, nx, ny, seed=53 + qid + 7 * coll)
    F = force_for(force, nx, ny)
    cm, taus = _cm(coll, qid, F)
    bcs = bc_dicts(bcname, nx, ny) if bcname != "periodic" else []
    q = _q(qid)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs))
    L.simulate(model, range(steps))
    S = f0.copy(order="F")
    C = f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, steps, force=F, bcs=bcs)
    _assert_same(model.f_stream.to_numpy(), S, f"first-step tile D2Q{qid} coll {coll} {bcname}")
    _assert_same(model.f_collision.to_numpy(), C, f"first-step tile D2Q{qid} coll {coll} {bcname} (C)")


@pytest.mark.parametrize("qid", [4, 5, 9])
@pytest.mark.parametrize("storage", ["f32", "f16dev", "bf16dev", "f16neq", "bf16neq"])
@pytest.mark.parametrize("coll", [0, 1])
@pytest.mark.parametrize("bcname,force", [("periodic", "none"), ("couette", "const"),
                                          ("cavity", "field"), ("partial", "none")])
def test_first_step_tile_reduced(oracle_lib, qid, storage, coll, bcname, force):
    """The reduced storages on lattices of whole 64 x 8 tiles start a run with
    one pass (the halo tile's collide! output kept in LDS as stored).  Against
    the two-pass start (collide!, then the fused step; the per-launch timing
    entry point keeps it): equal up to the rounding of the FP32 operations,
    which the kernels order differently -- for the 16-bit storages that flips
    single 16-bit roundings, as for the other kernel pairs; a wrong pull, halo
    or boundary value moves populations by ~1e-2 w.  Against the oracle's
    FP64 run: the same velocity error as the two-pass start (the storage's own
    error; BF16 on the D2Q4/5 sets of random populations reaches 6e-4)."""
    from oracle import lbm_numpy as ln
    L = _L()
    nx, ny = 128, 16
    steps = 6
    q = _q(qid)
    f0 = random_populations(qid, nx, ny, seed=61 + qid, noise=1e-4)
    F = force_for(force, nx, ny)
    cm, taus = _cm(coll, qid, F)
    bcs = bc_dicts(bcname, nx, ny) if bcname != "periodic" else []
    one = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs), storage=storage)
    L.simulate(one, range(steps))
    two = L.LatticeBoltzmannModel.from_populations(q, f0, cm, to_lbmhip(bcs), storage=storage)
    L.simulate_launch_times(two, steps)
    w = np.asarray(q.weights, dtype=float)
    ulp_tol = {"f32": 1e-5, "f16dev": 1e-4, "bf16dev": 1e-3, "f16neq": 1e-4,
               "bf16neq": 1e-3}[storage]
    for a, b in ((one.f_stream.to_numpy(), two.f_stream.to_numpy()),
                 (one.f_collision.to_numpy(), two.f_collision.to_numpy())):
        d = np.abs((a - b) / w)
        assert d.max() <= ulp_tol, d.max()
        assert np.count_nonzero(d > 1e-6) <= 0.1 * d.size
    S, C = f0.copy(order="F"), f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, steps, force=F, bcs=bcs)
    rho, ux, uy = ln.moments_field(ln.by_id(qid), S)

    def err(model):
        m = L.moments(model.f_stream, stress=False)
        num = np.sqrt(np.sum((ux - m.ux) ** 2 + (uy - m.uy) ** 2))
        return num / np.sqrt(np.sum(ux ** 2 + uy ** 2))

    e1, e2 = err(one), err(two)
    assert e2 < 4 * _LP_TOL[storage], e2
    assert e1 <= 1.05 * e2 + 1e-6, (e1, e2)
