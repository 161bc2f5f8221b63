"""Benchmark: D2Q9 TRT Taylor-Green vortex, periodic 4096 x 4096, on one MI355X.

A step is one time step of the lattice (collide -> stream -> BCs) over the
whole 4096^2 lattice, i.e. one fused pull collide-stream kernel launch.
``value`` is the plain FP64 throughput in MLUPS (million lattice updates per
second) over K timed steps, populations resident in HBM.  The same run with
each reduced storage (FP16/BF16 non-equilibrium records, FP16/BF16
deviations, shifted FP32; FP32 arithmetic) is reported beside it with its
relative L2 velocity error against the FP64 run.

``roofline`` prices the fused kernel against MI355X's 8 TB/s: algorithmic
bytes per launch = 2 * Q * sizeof(population) * nodes (one read and one write
of every population), divided by the kernel's average duration measured with
HIP events on the library's stream.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "latticeboltzmann.jl_amd"))

METRIC = ("MLUPS + %HBM roofline, D2Q9 TRT 4096²: FP16-store vs FP64 plain at "
          "err≤1e-4")
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "hbm_traffic.json")


def tgv_populations(L, q, n, u0=0.02):
    c = (np.arange(n) + 0.5) * (2 * np.pi / n)
    X, Y = np.meshgrid(c, c, indexing="ij")
    ux = u0 * np.sin(X) * np.cos(Y)
    uy = -u0 * np.cos(X) * np.sin(Y)
    rho = 1.0 - (3.0 * u0 * u0 / 4.0) * (np.cos(2 * X) + np.cos(2 * Y))
    return L.equilibrium_populations(q, rho, ux, uy)


# the Taylor-Green workload's length: the accuracy of every storage is taken
# after this many steps, whatever --steps is
ACCURACY_STEPS = 5000
# fused launches timed one by one for the spread of the kernel's duration
SPREAD_LAUNCHES = 200
# back-to-back device copies of the FP64 lattice for the copy-rate ceiling
COPY_LAUNCHES = 50


def run_storage(L, q, f0, cm, storage, steps, warmup, accuracy_steps):
    """An untimed ``accuracy_steps``-step run of this storage (its velocity is
    returned), then, on a second model uploaded before that run, W untimed
    warm-up steps and K timed steps (wall clock around them, HIP events around
    the fused launches), with the shader clock sampled before and after the
    timed steps.  The accuracy run goes first and its velocity is read back
    only after the timed steps, so the warm-up follows it on the device with
    no idle stretch in between: the timed launches start from the GPU's
    sustained state, not from the idle stretch of a host upload or download
    (clock-sensitive kernels run their first launches after idle up to 10%
    slower)."""
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm, storage=storage)
    acc = L.LatticeBoltzmannModel.from_populations(q, f0, cm, storage=storage)
    print(f"[bench] {storage}: accuracy run, {accuracy_steps} untimed steps", file=sys.stderr,
          flush=True)
    L.simulate(acc, range(accuracy_steps))
    print(f"[bench] {storage}: {warmup} warmup + {steps} timed steps", file=sys.stderr,
          flush=True)
    if warmup:
        L.simulate(model, range(warmup))
    L.synchronize()
    c0 = L.clock_sample()
    t0 = time.perf_counter()
    kernel_ms, nlaunch = L.simulate_timed(model, steps)
    L.synchronize()
    wall = time.perf_counter() - t0
    c1 = L.clock_sample()
    m = L.moments(acc.f_stream, stress=False)
    del acc
    return model, wall, kernel_ms, nlaunch, L.average_clock_mhz(c0, c1), (m.ux, m.uy)


def copy_rate(L, f, copies=COPY_LAUNCHES):
    """Rate of a plain device copy of the same bytes on the same GPU
    (``lbm_field_copy``: 16-byte lanes, non-temporal stores), ``copies``
    launches back to back timed by the host clock: the practical ceiling a
    kernel that reads and writes every byte once is read against."""
    g = f.copy()
    for _ in range(3):
        g.copyto_(f)
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(copies):
        g.copyto_(f)
    L.synchronize()
    dt = (time.perf_counter() - t0) / copies
    g.close()
    return dt


def launch_spread(L, model, launches=SPREAD_LAUNCHES):
    """Per-launch duration of the fused kernel over a run of ``launches``
    launches (untimed for the bench line; HIP events between launches), with
    the average shader clock over that run."""
    c0 = L.clock_sample()
    t = L.simulate_launch_times(model, launches + 1) * 1e3
    c1 = L.clock_sample()
    return {"launches": int(t.size), "us_min": round(float(t.min()), 2),
            "us_avg": round(float(t.mean()), 2), "us_max": round(float(t.max()), 2),
            "us_first_20_avg": round(float(t[:20].mean()), 2),
            "us_last_20_avg": round(float(t[-20:].mean()), 2),
            "shader_clock_mhz": _r(L.average_clock_mhz(c0, c1))}


def _r(v, nd=1):
    return None if v is None else round(v, nd)


def rel_l2(a, b):
    num = np.sqrt(np.sum((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2))
    return float(num / np.sqrt(np.sum(b[0] ** 2 + b[1] ** 2)))


def traffic_for(key):
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh).get(key)
    except (OSError, ValueError):
        return None


def host_cpus():
    """CPUs this process may run on: the affinity mask, capped by a cgroup
    CPU quota when one is set; with the host's logical CPU count and model."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = min(usable, quota) if quota else usable
    return {"threads": threads, "nproc": os.cpu_count(), "affinity": usable,
            "cgroup_quota_cpus": quota, "model": model}


def cpu_baseline(n, tau_s, tau_a, f0, target_s):
    """The FP64 CPU restatement (the oracle) timed on this host's cores on a
    bounded sample of the same workload: as many steps of the 4096^2 lattice
    as fit in about target_s seconds (at most 400), on every CPU the process
    may use, and the same on one core."""
    from oracle import clib
    host = host_cpus()
    threads = host["threads"]
    clib.set_threads(threads)
    S = np.asfortranarray(f0.copy())
    C = np.asfortranarray(f0.copy())
    t0 = time.perf_counter()
    clib.simulate(9, S, C, 1, [tau_s, tau_a], 1)
    t1 = time.perf_counter() - t0
    steps = max(1, min(400, int(target_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    clib.simulate(9, S, C, 1, [tau_s, tau_a], steps)
    dt = time.perf_counter() - t0
    mlups = n * n * steps / dt / 1e6
    # the reference's own benchmarks run one thread: the same restatement on one core
    clib.set_threads(1)
    t0 = time.perf_counter()
    clib.simulate(9, S, C, 1, [tau_s, tau_a], 1)
    t1 = time.perf_counter() - t0
    steps1 = max(1, min(50, int(0.4 * target_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    clib.simulate(9, S, C, 1, [tau_s, tau_a], steps1)
    dt1 = time.perf_counter() - t0
    clib.set_threads(threads)
    return {"value": round(mlups, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"D2Q9 TRT Taylor-Green {n}x{n}, FP64, {steps} steps "
                      f"({dt:.1f} s) with the C restatement on {threads} OpenMP threads",
            "host": host,
            "single_core": {"value": round(n * n * steps1 / dt1 / 1e6, 2), "unit": "MLUPS",
                            "sample": f"{steps1} steps ({dt1:.1f} s) on 1 thread"}}


def processing_leg(L, q, model, n, u0, tau_s, t_steps, calls=20):
    """One row of TrackHydrodynamicErrors on the FP64 4096^2 state, on the
    device (lbm_track_errors: one pass over the populations and eight expected
    planes, 11 numbers back) against the host path the reference takes
    (download the populations, then the oracle's node loop on one core)."""
    c = (np.arange(n) + 0.5) * (2 * np.pi / n)
    X, Y = np.meshgrid(c, c, indexing="ij")
    k = 2 * np.pi / n
    nu = (tau_s - 0.5) / q.speed_of_sound_squared
    # the vortex of tgv_populations in units of u0, its density = pressure, and
    # -nu (grad u + grad u^T) / u0 as the stress
    ux, uy = np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y)
    prho = -(3.0 * u0 * u0 / 4.0) * (np.cos(2 * X) + np.cos(2 * Y))
    sxx = -nu * 2 * k * np.cos(X) * np.cos(Y) / u0
    sxy = np.zeros_like(sxx)
    planes = np.stack([prho, ux, uy, prho, sxx, sxy, sxy, -sxx], axis=-1)
    d = float(np.exp(-2 * nu * k * k * t_steps))
    affine = (1.0, d * d, 0.0, d, 0.0, d, 1.0, d * d, 0.0, d, 0.0, d, 0.0, d, 0.0, d)
    ex = L.ExpectedFields(n, n, planes)
    tau_p = q.speed_of_sound_squared * nu
    row = L.track_errors(model.f_stream, ex, tau_p, u0, 1.0, affine)
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        row = L.track_errors(model.f_stream, ex, tau_p, u0, 1.0, affine)
    dev_ms = (time.perf_counter() - t0) / calls * 1e3
    per_node = q.Q * 8 + 8 * 8
    out = {
        "workload": f"one TrackHydrodynamicErrors row, D2Q9 FP64 {n}x{n} after {t_steps} steps",
        "device_ms_per_row": round(dev_ms, 3),
        "bytes_per_row": per_node * n * n,
        "device_GBps_wall": round(per_node * n * n / (dev_ms * 1e-3) / 1e9, 1),
        "error_rho": float(row[0]), "error_u": float(row[1]), "error_p": float(row[2]),
        "mass": float(row[7]), "energy": float(row[9]),
    }
    from oracle import clib
    t0 = time.perf_counter()
    f = model.f_stream.to_numpy()
    down_ms = (time.perf_counter() - t0) * 1e3
    clib.set_threads(1)
    t0 = time.perf_counter()
    ref = clib.track_errors(9, f, np.asfortranarray(planes), affine, tau_p, u0, 1.0)
    host_ms = (time.perf_counter() - t0) * 1e3
    del f
    out.update({
        "host_path_ms": round(down_ms + host_ms, 1),
        "host_path": "download the populations + the oracle's node loop on 1 core",
        "download_ms": round(down_ms, 1),
        "speedup_vs_host_path": round((down_ms + host_ms) / dev_ms, 1),
        # the norms, mass and energy against the oracle's sequential sums
        # (error_sigma_xy is 0/0 for this vortex, whose analytic sigma_xy
        # vanishes; the momentum cancels to rounding, compared against the mass)
        "max_rel_diff_vs_oracle": float(np.max([abs(a - b) / max(abs(b), 1e-300)
                                                for k, (a, b) in enumerate(zip(row[:10], ref[:10]))
                                                if np.isfinite(b) and k != 8])),
        "momentum_diff_over_mass": float(abs(row[8] - ref[8]) / abs(ref[7])),
    })
    return out


def shear_wave_populations(L, q, n, amp, k):
    y = np.arange(n)
    ux = np.tile(amp * np.sin(2 * np.pi * k * y / n), (n, 1))
    return L.equilibrium_populations(q, np.ones((n, n)), ux, np.zeros((n, n)))


def batch_leg(L, steps=100, count=512, seed=0):
    """512 independent periodic D2Q9 TRT lattices (sizes log-uniform in
    32^2..1024^2, tau ~ U(0.52, 1.5), shear waves of amplitude U(0.001, 0.05)
    and wavenumber 1, 2 or 4; NumPy PCG64 seed 0), FP32 storage: one batched
    launch per step with 100-step hipGraph chunks, against the same lattices
    advanced one after another with per-lattice launches."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(32), np.log(1024), count))).astype(int)
    tau = rng.uniform(0.52, 1.5, count)
    amp = rng.uniform(0.001, 0.05, count)
    kk = rng.choice([1, 2, 4], count)
    tau_a = 0.5 + 0.25 / (tau - 0.5)
    q = L.D2Q9()
    b = L.LatticeBatch(q, [(n, n) for n in sizes], tau, tau_a, storage="f32")
    for k in range(count):
        b.upload(k, shear_wave_populations(L, q, int(sizes[k]), amp[k], kk[k]))
    tuned = b.autotune(3)
    b.simulate(steps, graph_chunk=steps)                      # warm-up (builds the graph)
    L.synchronize()
    t0 = time.perf_counter()
    fused_ms = b.simulate(steps, graph_chunk=steps)
    wall_b = time.perf_counter() - t0
    nodes = int(np.sum(sizes.astype(np.int64) ** 2))
    seq = 0.0
    for k in range(count):
        f0 = shear_wave_populations(L, q, int(sizes[k]), amp[k], kk[k])
        m = L.LatticeBoltzmannModel.from_populations(q, f0, L.TRT.fields(tau[k], tau_a[k]),
                                                     storage="f32")
        L.simulate(m, range(2))
        L.synchronize()
        t0 = time.perf_counter()
        L.simulate(m, range(steps))
        L.synchronize()
        seq += time.perf_counter() - t0
        del m
    del b
    return {"lattices": count, "nodes": nodes, "steps": steps, "storage": "f32",
            "graph_chunk": steps, "launch": tuned,
            "batched_mlups": round(nodes * steps / wall_b / 1e6, 1),
            "batched_fused_kernel_ms": round(fused_ms, 3),
            "sequential_mlups": round(nodes * steps / seq / 1e6, 1),
            "speedup_vs_sequential": round(seq / wall_b, 2)}


def adjoint_leg(L, n=512, T=2000, tau=0.7):
    """Reverse-mode gradients of L = 1/2 |u(T) - u_target|^2 for D2Q9 SRT on a
    periodic n x n lattice: u0 and u_target are divergence-free fields from
    Fourier modes |k| <= 8 at amplitude 0.02 (seeds 1 and 2), sqrt(T)
    checkpointing.  FP64 gradients are checked against h = 1e-4 central
    differences (tau, and one divergence-free direction in u0); FP32 storage
    gradients against FP64.  One gradient = forward pass + recomputation +
    reverse pass, 3 T lattice sweeps."""
    q = L.D2Q9()
    u0 = L.divergence_free_field((n, n), 8, 0.02, seed=1)
    ut = L.divergence_free_field((n, n), 8, 0.02, seed=2)
    out = {"workload": f"D2Q9 SRT tau={tau}, periodic {n}x{n}, T={T}, loss 1/2|u(T)-u_target|^2"}
    res = {}
    for st in ("f64", "f32"):
        prob = L.AdjointProblem(q, L.SRT(tau), (n, n), T, storage=st)
        prob.gradient(u0, ut)                                   # warm-up
        L.synchronize()
        t0 = time.perf_counter()
        r = prob.gradient(u0, ut)
        wall = time.perf_counter() - t0
        res[st] = r
        info = prob.info()
        out[st] = {"gradient_ms": round(wall * 1e3, 2), "device_ms": round(r["ms"], 2),
                   "checkpoint_every": info["checkpoint_every"],
                   "device_bytes": info["bytes"], "loss": r["loss"],
                   "grad_tau": r["grad_tau"],
                   "sweeps_mlups": round(3 * n * n * T / wall / 1e6, 1)}
        if st == "f64":
            h = 1e-4
            prob.set_collision(L.SRT(tau + h))
            lp = prob.loss(u0, ut)
            prob.set_collision(L.SRT(tau - h))
            lm = prob.loss(u0, ut)
            prob.set_collision(L.SRT(tau))
            fd_t = (lp - lm) / (2 * h)
            v = L.divergence_free_field((n, n), 8, 0.02, seed=3)
            fd_u = (prob.loss(u0 + h * v, ut) - prob.loss(u0 - h * v, ut)) / (2 * h)
            ad_u = float(np.sum(r["grad_u0"] * v))
            out[st]["fd_rel_err_tau"] = abs(r["grad_tau"] - fd_t) / abs(fd_t)
            out[st]["fd_rel_err_u0_direction"] = abs(ad_u - fd_u) / abs(fd_u)
        prob.close()
    a, b = res["f32"], res["f64"]
    out["f32_vs_f64"] = {
        "grad_u0_rel_l2": float(np.linalg.norm(a["grad_u0"] - b["grad_u0"])
                                / np.linalg.norm(b["grad_u0"])),
        "grad_tau_rel": abs(a["grad_tau"] / b["grad_tau"] - 1.0),
        "loss_rel": abs(a["loss"] / b["loss"] - 1.0)}
    return out


def startup_profile(flow, ny, nu, t, F=1e-6, u_wall=0.05, nterms=4000):
    """Analytic start-up of plane Poiseuille (body force F, walls at rest) and
    Couette (top wall u_wall) flow from rest between halfway walls at y = 0 and
    y = ny (node centres y_j = j + 1/2), after t steps of viscosity nu:
    Fourier sine series of the diffusion equation."""
    h = float(ny)
    y = np.arange(ny) + 0.5
    n = np.arange(1, nterms + 1, dtype=np.float64)
    decay = np.exp(-nu * (n * np.pi / h) ** 2 * t)
    modes = np.sin(np.outer(n, np.pi * y / h))
    if flow == "poiseuille":
        coef = np.where(n % 2 == 1, 4.0 * F * h * h / (nu * (n * np.pi) ** 3), 0.0)
        return (coef * (1.0 - decay)) @ modes
    coef = 2.0 * u_wall * (-1.0) ** n / (n * np.pi)
    return u_wall * y / h + (coef * decay) @ modes


def mrt_leg(L, steps=20):
    """Regularised MRT (mrt.jl:56-118) in FP64 on the periodic Taylor-Green
    vortex, D2Q9 at 4096^2 and D2Q17 / D2Q37 at 8192 x 2048, all relaxation
    times 0.8, beside SRT(0.8) on the same lattice: kernel time per step from
    HIP events and the fraction of 8 TB/s at 2 Q 8 bytes per node."""
    rows = []
    for qid, nx, ny in ((9, 4096, 4096), (17, 8192, 2048), (37, 8192, 2048)):
        q = L.Quadratures[f"D2Q{qid}"]()
        x = (np.arange(nx) + 0.5) * (2 * np.pi / nx)
        y = (np.arange(ny) + 0.5) * (2 * np.pi / ny)
        X, Y = np.meshgrid(x, y, indexing="ij")
        f0 = L.equilibrium_populations(q, np.ones((nx, ny)), 0.02 * np.sin(X) * np.cos(Y),
                                       -0.02 * np.cos(X) * np.sin(Y))
        del X, Y
        N = max(2, L.order(q) // 2)
        # "MRT-MFMA": the same collision on the matrix cores (16-node tiles)
        for name, cm in (("MRT", L.MRT(q, [0.8] * N)), ("MRT-MFMA", L.MRT(q, [0.8] * N)),
                         ("SRT", L.SRT(0.8))):
            print(f"[bench] D2Q{qid} {nx}x{ny} FP64 {name}: {steps} steps", file=sys.stderr,
                  flush=True)
            m = L.LatticeBoltzmannModel.from_populations(q, f0, cm)
            if name == "MRT-MFMA":
                m.launch = (256, 1, 16, 1)
            L.simulate(m, range(3))
            ms, nl = L.simulate_timed(m, steps)
            us = ms / max(nl, 1) * 1e3
            gbs = 2 * q.Q * 8 * nx * ny / (us * 1e-6) / 1e9
            rows.append({"quadrature": f"D2Q{qid}", "lattice": [nx, ny], "collision": name,
                         "us_per_step": round(us, 1), "mlups": round(nx * ny / us, 1),
                         "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
            del m
        del f0
    return rows


def sizes_leg(L, sizes, steps=20):
    """D2Q9 TRT(0.8, 1.1) on periodic n x n lattices (uniform flow 0.01 along
    x, equilibrium start) in FP64 and FP32: kernel time per step from HIP
    events and algorithmic GB/s (2 Q sizeof per node).  Rows of whole 128-byte
    lines run the row-shaped kernel, other widths the flat-index one."""
    q = L.D2Q9()
    rows = []
    for n in sizes:
        f0 = L.equilibrium_populations(q, np.ones((n, n)), np.full((n, n), 0.01),
                                       np.zeros((n, n)))
        for storage, eb in (("f64", 8), ("f32", 4)):
            print(f"[bench] D2Q9 {n}x{n} {storage}: {steps} steps", file=sys.stderr, flush=True)
            m = L.LatticeBoltzmannModel.from_populations(q, f0, L.TRT.fields(0.8, 1.1),
                                                         storage=storage)
            L.simulate(m, range(3))
            ms, nl = L.simulate_timed(m, steps)
            us = ms / max(nl, 1) * 1e3
            gbs = 2 * q.Q * eb * n * n / (us * 1e-6) / 1e9
            rows.append({"lattice": [n, n], "storage": storage,
                         "kernel": "row-shaped" if (n * eb) % 128 == 0 else "flat-index",
                         "us_per_step": round(us, 1), "achieved_GBps": round(gbs, 1),
                         "frac": round(gbs / HBM_PEAK_GBS, 4)})
            del m
        del f0
    return rows


def precision_leg(L, steps, nx=8192, ny=2048, tau=0.8, qids=(9, 17, 37),
                  flows=("poiseuille", "couette"),
                  storages=("f64", "f32", "f16dev", "bf16dev", "f16neq", "bf16neq"), force=1e-6):
    """D2Q9 / D2Q17 / D2Q37 SRT(tau) on an nx x ny channel (periodic in x,
    walls at the bottom and top rows), from rest: Poiseuille with the constant
    body force (1e-6, 0) and bounce-back walls (poiseuille.jl:93-96), Couette
    with a bounce-back bottom and a top wall moving at (0.05, 0)
    (couette_flow.jl:82-85).  Each storage runs the same steps; the velocity
    field is compared with the FP64 run (relative L2 over the lattice) and with
    the analytic start-up profile."""
    rows = []
    for qid in qids:
        q = L.Quadratures[f"D2Q{qid}"]()
        nu = (tau - 0.5) / q.speed_of_sound_squared
        f0 = L.equilibrium_populations(q, np.ones((nx, ny)), np.zeros((nx, ny)),
                                       np.zeros((nx, ny)))
        for flow in flows:
            full = ((1, nx), (1, ny))
            if flow == "poiseuille":
                cm = L.SRT(tau, np.array([force, 0.0]))
                bcs = [L.BounceBack(L.North(), *full), L.BounceBack(L.South(), *full)]
            else:
                cm = L.SRT(tau)
                bcs = [L.BounceBack(L.South(), *full),
                       L.MovingWall(L.North(), *full, (0.05, 0.0))]
            ua = startup_profile(flow, ny, nu, steps, F=force)
            ref = None
            for st in storages:
                m = L.LatticeBoltzmannModel.from_populations(q, f0, cm, bcs, storage=st)
                L.synchronize()
                t0 = time.perf_counter()
                kms, nl = L.simulate_timed(m, steps)
                L.synchronize()
                wall = time.perf_counter() - t0
                mom = L.moments(m.f_stream, stress=False)
                ux, uy = mom.ux, mom.uy
                if st == "f64":
                    ref = (ux, uy)
                den = np.sqrt(np.sum(ref[0] ** 2 + ref[1] ** 2))
                e64 = float(np.sqrt(np.sum((ux - ref[0]) ** 2 + (uy - ref[1]) ** 2)) / den)
                ea = float(np.sqrt(np.sum((ux - ua[None, :]) ** 2 + uy ** 2))
                           / np.sqrt(nx * np.sum(ua ** 2)))
                row = {"quadrature": f"D2Q{qid}", "flow": flow, "storage": st, "steps": steps,
                       "rel_l2_vs_f64": e64, "rel_l2_vs_analytic": ea,
                       "mlups": round(nx * ny * steps / wall / 1e6, 1),
                       "fused_kernel_ms": round(kms / max(nl, 1), 5),
                       "u_max": float(np.abs(ux).max())}
                rows.append(row)
                print(json.dumps(row), flush=True)
                del m
        del f0
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the Taylor-Green workload: 5000 timed steps after 50 warm-up steps
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--tau", type=float, default=0.6)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--accuracy-steps", type=int, default=ACCURACY_STEPS,
                    help="untimed steps of the FP64-vs-storage velocity comparison (the "
                         "workload's 5000; fewer only for profiler runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-processing", action="store_true")
    ap.add_argument("--precision", action="store_true",
                    help="run only the 8192x2048 channel precision table (minutes)")
    ap.add_argument("--precision-steps", type=int, default=20000)
    ap.add_argument("--precision-quadratures", default="9,17,37")
    ap.add_argument("--precision-storages", default="f64,f32,f16dev,bf16dev,f16neq,bf16neq",
                    help="storages of the channel table (f64 first: the reference run)")
    ap.add_argument("--precision-force", type=float, default=1e-6,
                    help="body force of the Poiseuille channel")
    ap.add_argument("--precision-flows", default="poiseuille,couette")
    ap.add_argument("--mrt", action="store_true",
                    help="run only the FP64 MRT-against-SRT table (D2Q9/17/37)")
    ap.add_argument("--sizes", default="",
                    help="run only D2Q9 TRT FP64/FP32 on these n x n lattices (e.g. 4096,4100)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: liblbm_hip runs on one GPU; --gpus {a.gpus} is not supported "
              "(sharding the lattice as y-slabs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    if a.accuracy_steps < 1:
        ap.error("--accuracy-steps must be at least 1")

    import lbmhip as L
    if L.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.mrt:
        print(json.dumps({"mrt": mrt_leg(L)}), flush=True)
        return
    if a.sizes:
        sizes = [int(v) for v in a.sizes.split(",")]
        print(json.dumps({"sizes": sizes_leg(L, sizes)}), flush=True)
        return
    if a.precision:
        qids = tuple(int(v) for v in a.precision_quadratures.split(","))
        rows = precision_leg(L, a.precision_steps, qids=qids,
                             storages=tuple(a.precision_storages.split(",")),
                             force=a.precision_force,
                             flows=tuple(a.precision_flows.split(",")))
        print(json.dumps({"precision": rows}), flush=True)
        return
    n, K, W = a.n, a.steps, a.warmup
    q = L.D2Q9()
    cm = L.trt_magic(a.tau, 0.25)
    print(f"[bench] building Taylor-Green {n}x{n} initial populations", file=sys.stderr,
          flush=True)
    f0 = tgv_populations(L, q, n)
    nodes = n * n

    acc_steps = a.accuracy_steps
    m64, wall64, kms64, nl64, clk64, u64 = run_storage(L, q, f0, cm, "f64", K, W, acc_steps)
    mlups64 = nodes * K / wall64 / 1e6
    bytes64 = 2 * q.Q * 8 * nodes
    kavg64 = kms64 / max(nl64, 1)
    gbs64 = bytes64 / (kavg64 * 1e-3) / 1e9
    spread64 = launch_spread(L, m64)
    copy_s = copy_rate(L, m64.f_stream)
    copy_gbs = bytes64 / copy_s / 1e9

    result = {
        "metric": METRIC,
        "value": round(mlups64, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall64 / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"D2Q9 TRT (magic Lambda=1/4, tau={a.tau}) Taylor-Green vortex, "
                        f"periodic {n}x{n}, plain FP64 populations, fused pull collide-stream",
            "lattice": [n, n],
            "quadrature": "D2Q9",
            "collision": "TRT",
            "tau_symmetric": cm.τ_symmetric,
            "tau_asymmetric": cm.τ_asymmetric,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs64, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs64 / HBM_PEAK_GBS, 4),
            "traffic": traffic_for("d2q9_trt_f64_4096"),
            "kernel": "k_fused_v<9,TRT,F64> (128x2 blocks, 1 node/thread, non-temporal stores)",
            "kernel_ms_avg": round(kavg64, 5),
            "bytes_per_launch": bytes64,
            "shader_clock_mhz": _r(clk64),
            "spread": spread64,
            "copy": {"gbs": round(copy_gbs, 1), "us_per_copy": round(copy_s * 1e6, 2),
                     "frac_of_copy": round(gbs64 / copy_gbs, 4),
                     "how": f"lbm_field_copy of the FP64 lattice ({bytes64 // 2} bytes, "
                            f"read and written; 16-byte lanes, non-temporal stores), "
                            f"{COPY_LAUNCHES} launches back to back, host clock"},
        },
    }

    if not a.no_fp16:
        # the reduced storages: non-equilibrium parts about a 16-bit equilibrium
        # record (*neq: 2 Q + 8 bytes per node), rest-state deviations (*dev),
        # and shifted FP32
        kernels = {
            "f16neq": "k_fused_neq_rows2<9,TRT,F16Neq> (two columns x 2 rows per lane, LDS "
                      "record window, edge nodes in its last block column)",
            "bf16neq": "k_fused_neq_rows2<9,TRT,BF16Neq> (two columns x 2 rows per lane, LDS "
                       "record window, edge nodes in its last block column)",
            "f16dev": "k_fused_v<9,TRT,F16Dev> (4 nodes/thread)",
            "bf16dev": "k_fused_v<9,TRT,BF16Dev> (4 nodes/thread)",
            "f32": "k_fused_v<9,TRT,F32> (256x1 blocks, 1 node/thread)",
        }
        for storage, key, per_node in (("f16neq", "fp16neq", 2 * q.Q + 8),
                                       ("f16dev", "fp16dev", 2 * q.Q),
                                       ("bf16neq", "bf16neq", 2 * q.Q + 8),
                                       ("bf16dev", "bf16dev", 2 * q.Q),
                                       ("f32", "fp32", 4 * q.Q)):
            m16, wall16, kms16, nl16, clk16, u16 = run_storage(L, q, f0, cm, storage, K, W,
                                                               acc_steps)
            mlups16 = nodes * K / wall16 / 1e6
            bytes16 = 2 * per_node * nodes
            kavg16 = kms16 / max(nl16, 1)
            gbs16 = bytes16 / (kavg16 * 1e-3) / 1e9
            spread16 = launch_spread(L, m16)
            del m16
            result[key] = {
                "value": round(mlups16, 1),
                "unit": "MLUPS",
                "storage": storage,
                "ratio_vs_f64": round(mlups16 / mlups64, 3),
                "rel_l2_velocity_err_vs_f64": rel_l2(u16, u64),
                "steps_compared": acc_steps,
                "ms_per_step": round(wall16 / K * 1e3, 5),
                "roofline": {"bound": "hbm", "achieved": round(gbs16, 1), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(gbs16 / HBM_PEAK_GBS, 4),
                             "traffic": traffic_for(f"d2q9_trt_{storage}_4096"),
                             "kernel": kernels[storage],
                             "kernel_ms_avg": round(kavg16, 5), "bytes_per_launch": bytes16,
                             "shader_clock_mhz": _r(clk16), "spread": spread16},
            }

    if not a.no_processing:
        print("[bench] on-device processing row", file=sys.stderr, flush=True)
        result["processing"] = processing_leg(L, q, m64, n, 0.02, cm.τ_symmetric,
                                              W + K + SPREAD_LAUNCHES + 1)
    del m64
    if not a.no_batch:
        print("[bench] batch of 512 lattices", file=sys.stderr, flush=True)
        result["batch"] = batch_leg(L)
    if not a.no_adjoint:
        print("[bench] adjoint gradients 512^2, T=2000", file=sys.stderr, flush=True)
        result["adjoint"] = adjoint_leg(L)
    if not a.no_cpu:
        print("[bench] CPU baseline (oracle restatement)", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(n, cm.τ_symmetric, cm.τ_asymmetric, f0,
                                              a.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
