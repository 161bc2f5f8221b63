"""GPU reverse pass (lbm_adjoint_*) against the oracle and finite differences.

* forward: u(S_T) of the adjoint's own forward run is bit-exact with the C
  restatement's FP64 run from the same equilibrium start;
* gradients: dL/du0 node by node and dL/dtau against the reverse-mode checker
  (oracle/lbm_adjoint_numpy.py), every quadrature, SRT and TRT;
* central finite differences (h = 1e-4) of the FP64 run, on the GPU forward and
  on the C restatement's forward, along random directions and along tau:
  relative agreement 1e-6;
* checkpoint interval: any interval gives the same bits (recomputation is
  deterministic, the reductions have a fixed order);
* the differentiation workload: D2Q9 SRT, periodic 512 x 512, tau = 0.7,
  T = 2000, divergence-free u0 from modes |k| <= 8 at amplitude 0.02: FP64
  gradients against finite differences, FP32 storage gradients against FP64.
"""
import numpy as np
import pytest

from lbm_cases import QIDS
from oracle import lbm_adjoint_numpy as la
from oracle import lbm_numpy as ln

pytestmark = pytest.mark.gpu


def _L():
    import lbmhip
    return lbmhip


def _q(qid):
    return _L().Quadratures[f"D2Q{qid}"]()


def _cm(coll, taus):
    L = _L()
    return L.SRT(taus[0]) if coll == 0 else L.TRT.fields(taus[0], taus[1])


def _fields(nx, ny, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    return amp * rng.standard_normal((nx, ny, 2)), amp * rng.standard_normal((nx, ny, 2))


def _oracle_uT(oracle_lib, qid, u0, coll, taus, T, rho0=1.0):
    nx, ny = u0.shape[:2]
    f = np.empty((nx, ny, ln.by_id(qid).Q), order="F")
    for x in range(nx):
        for y in range(ny):
            f[x, y] = oracle_lib.equilibrium(qid, rho0, u0[x, y])
    S, C = f.copy(order="F"), f.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, T)
    _, ux, uy = oracle_lib.moments(qid, S, taus[0])[:3]
    return np.stack([ux, uy], axis=-1)


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1])
def test_forward_bit_exact(oracle_lib, qid, coll):
    L = _L()
    nx, ny, T = 19, 14, 23
    taus = [0.8] if coll == 0 else [0.7, 1.3]
    u0, ut = _fields(nx, ny, seed=qid + coll)
    prob = L.AdjointProblem(_q(qid), _cm(coll, taus), (nx, ny), T, checkpoint_every=5)
    loss, uT = prob.loss(u0, ut, return_velocity=True)
    ref = _oracle_uT(oracle_lib, qid, u0, coll, taus, T)
    assert np.array_equal(uT, ref), np.abs(uT - ref).max()
    assert abs(loss / (0.5 * np.sum((ref - ut) ** 2)) - 1.0) < 1e-13
    prob.close()


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1])
def test_gradient_matches_checker(qid, coll):
    L = _L()
    nx, ny, T = 11, 9, 17
    taus = [0.8] if coll == 0 else [0.75, 1.2]
    u0, ut = _fields(nx, ny, seed=10 + qid + coll)
    prob = L.AdjointProblem(_q(qid), _cm(coll, taus), (nx, ny), T, checkpoint_every=4)
    r = prob.gradient(u0, ut)
    Lref, gref, gtref = la.gradient(ln.by_id(qid), u0, ut, taus, coll, T)
    assert abs(r["loss"] / Lref - 1.0) < 1e-12
    err = np.abs(r["grad_u0"] - gref).max() / np.abs(gref).max()
    assert err < 1e-10, err
    gt = np.atleast_1d(r["grad_tau"])
    np.testing.assert_allclose(gt, gtref, rtol=1e-9, atol=1e-14)
    prob.close()


@pytest.mark.parametrize("coll", [0, 1])
def test_gradient_central_differences(oracle_lib, coll):
    """h = 1e-4 central differences of the GPU FP64 forward and of the C
    restatement's forward: both within 1e-6 relative of the reverse pass."""
    L = _L()
    nx, ny, T = 48, 40, 120
    taus = [0.7] if coll == 0 else [0.7, 1.1]
    u0 = L.divergence_free_field((nx, ny), 4, 0.02, seed=1)
    ut = L.divergence_free_field((nx, ny), 4, 0.02, seed=2)
    prob = L.AdjointProblem(_q(9), _cm(coll, taus), (nx, ny), T)
    r = prob.gradient(u0, ut)
    h = 1e-4
    v = np.random.default_rng(3).standard_normal(u0.shape)
    ad = float(np.sum(r["grad_u0"] * v))
    fd_gpu = (prob.loss(u0 + h * v, ut) - prob.loss(u0 - h * v, ut)) / (2 * h)
    assert abs(ad - fd_gpu) <= 1e-6 * abs(fd_gpu), (ad, fd_gpu)

    def c_loss(u, ts):
        uT = _oracle_uT(oracle_lib, 9, u, coll, ts, T)
        return 0.5 * float(np.sum((uT - ut) ** 2))

    fd_ref = (c_loss(u0 + h * v, taus) - c_loss(u0 - h * v, taus)) / (2 * h)
    assert abs(ad - fd_ref) <= 1e-6 * abs(fd_ref), (ad, fd_ref)
    gt = np.atleast_1d(r["grad_tau"])
    for k in range(len(taus)):
        tp, tm = list(taus), list(taus)
        tp[k] += h
        tm[k] -= h
        prob.set_collision(_cm(coll, tp))
        lp = prob.loss(u0, ut)
        prob.set_collision(_cm(coll, tm))
        lm = prob.loss(u0, ut)
        fd = (lp - lm) / (2 * h)
        assert abs(gt[k] - fd) <= 1e-6 * abs(fd), (k, gt[k], fd)
        fd_ref = (c_loss(u0, tp) - c_loss(u0, tm)) / (2 * h)
        assert abs(gt[k] - fd_ref) <= 1e-6 * abs(fd_ref), (k, gt[k], fd_ref)
    prob.set_collision(_cm(coll, taus))
    prob.close()


def test_checkpoint_interval_is_bitwise_invariant():
    L = _L()
    nx, ny, T = 33, 17, 50
    u0, ut = _fields(nx, ny, seed=5)
    outs = []
    for K in (1, 3, 7, 0, 50, 64):
        prob = L.AdjointProblem(_q(9), L.TRT.fields(0.65, 1.4), (nx, ny), T, checkpoint_every=K)
        info = prob.info()
        assert info["checkpoint_every"] == (K if K else 8)
        outs.append(prob.gradient(u0, ut))
        prob.close()
    for r in outs[1:]:
        assert r["loss"] == outs[0]["loss"]
        assert np.array_equal(r["grad_u0"], outs[0]["grad_u0"])
        assert r["grad_tau"] == outs[0]["grad_tau"]


@pytest.mark.parametrize("T", [0, 1, 2])
def test_short_runs(T):
    L = _L()
    nx, ny = 6, 5
    u0, ut = _fields(nx, ny, seed=6)
    prob = L.AdjointProblem(_q(9), L.SRT(0.9), (nx, ny), T)
    r = prob.gradient(u0, ut)
    Lref, gref, gtref = la.gradient(ln.by_id(9), u0, ut, [0.9], 0, T)
    np.testing.assert_allclose(r["grad_u0"], gref, rtol=0, atol=1e-13)
    assert abs(r["grad_tau"] - gtref[0]) <= 1e-12 * max(1.0, abs(gtref[0]))
    if T == 0:
        np.testing.assert_allclose(r["grad_u0"], u0 - ut, rtol=0, atol=1e-15)
        assert r["grad_tau"] == 0.0
    prob.close()


def test_f32_storage_tracks_f64():
    L = _L()
    n, T = 64, 300
    u0 = L.divergence_free_field((n, n), 6, 0.02, seed=11)
    ut = L.divergence_free_field((n, n), 6, 0.02, seed=12)
    res = {}
    for st in ("f64", "f32"):
        prob = L.AdjointProblem(_q(9), L.SRT(0.7), (n, n), T, storage=st)
        res[st] = prob.gradient(u0, ut)
        prob.close()
    a, b = res["f32"], res["f64"]
    rel_u = np.linalg.norm(a["grad_u0"] - b["grad_u0"]) / np.linalg.norm(b["grad_u0"])
    rel_t = abs(a["grad_tau"] / b["grad_tau"] - 1.0)
    rel_l = abs(a["loss"] / b["loss"] - 1.0)
    print(f"f32 vs f64: grad_u0 {rel_u:.3e}, grad_tau {rel_t:.3e}, loss {rel_l:.3e}")
    assert rel_u < 1e-4 and rel_t < 1e-4 and rel_l < 1e-5


def test_rrule_pullback():
    L = _L()
    q = _q(9)
    nx, ny, T = 12, 10, 15
    u0, ut = _fields(nx, ny, seed=8)
    y, back = L.rrule(L.simulate_loss, q, u0, 0.75, ut, T)
    assert y == L.simulate_loss(q, u0, 0.75, ut, T)
    t1 = back(1.0)
    t3 = back(-3.0)
    assert t1[0] is None and t1[1] is None
    np.testing.assert_allclose(t3[2], -3.0 * t1[2], rtol=1e-15)
    assert t3[3] == -3.0 * t1[3]
    y2, back2 = L.rrule(L.simulate_loss, q, u0, (0.75, 1.25), ut, T)
    assert len(back2(1.0)[3]) == 2


def test_adjoint_rejects_unsupported(oracle_lib):
    L = _L()
    q = _q(9)
    with pytest.raises(TypeError):
        L.AdjointProblem(q, L.MRT(q, [0.8, 0.9]), (8, 8), 10)
    with pytest.raises(ValueError):
        L.AdjointProblem(q, L.SRT(0.8), (8, 8), 10, storage="f16dev")
    with pytest.raises(L.LbmError):
        L.AdjointProblem(q, L.SRT(0.8), (8, 8), -1)
    prob = L.AdjointProblem(q, L.SRT(0.8), (8, 8), 10)
    with pytest.raises(ValueError):
        prob.gradient(np.zeros((8, 7, 2)), np.zeros((8, 8, 2)))
    prob.close()


def test_c5_workload_512():
    """The differentiation workload at full size, FP64 against central
    differences (h = 1e-4) and FP32 storage against FP64."""
    L = _L()
    n, T, tau = 512, 2000, 0.7
    u0 = L.divergence_free_field((n, n), 8, 0.02, seed=1)
    ut = L.divergence_free_field((n, n), 8, 0.02, seed=2)
    prob = L.AdjointProblem(_q(9), L.SRT(tau), (n, n), T)
    assert prob.info()["checkpoint_every"] == 45
    r = prob.gradient(u0, ut)
    h = 1e-4
    prob.set_collision(L.SRT(tau + h))
    lp = prob.loss(u0, ut)
    prob.set_collision(L.SRT(tau - h))
    lm = prob.loss(u0, ut)
    prob.set_collision(L.SRT(tau))
    fd_tau = (lp - lm) / (2 * h)
    v = L.divergence_free_field((n, n), 8, 0.02, seed=3)
    fd_u = (prob.loss(u0 + h * v, ut) - prob.loss(u0 - h * v, ut)) / (2 * h)
    ad_u = float(np.sum(r["grad_u0"] * v))
    print(f"C5 fp64: dL/dtau {r['grad_tau']:.10e} fd {fd_tau:.10e}; "
          f"<g,v> {ad_u:.10e} fd {fd_u:.10e}; {r['ms']:.1f} ms")
    assert abs(r["grad_tau"] - fd_tau) <= 1e-6 * abs(fd_tau)
    assert abs(ad_u - fd_u) <= 1e-6 * abs(fd_u)
    prob.close()
    p32 = L.AdjointProblem(_q(9), L.SRT(tau), (n, n), T, storage="f32")
    r32 = p32.gradient(u0, ut)
    p32.close()
    rel_u = np.linalg.norm(r32["grad_u0"] - r["grad_u0"]) / np.linalg.norm(r["grad_u0"])
    rel_t = abs(r32["grad_tau"] / r["grad_tau"] - 1.0)
    print(f"C5 fp32 vs fp64: grad_u0 {rel_u:.3e}, grad_tau {rel_t:.3e}; {r32['ms']:.1f} ms")
    assert rel_u < 1e-3 and rel_t < 1e-3
