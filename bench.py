"""Headline benchmark: checkerboard Metropolis sweeps of a 256^3 Lebwohl-Lasher
bulk lattice on one MI355X.

A step is one full sweep (both colours) of the 256^3 periodic lattice.  The K
timed sweeps are split evenly over the temperatures 1.2, 1.13, 1.0 and 0.6
(the thermalGPU workload), with the lattice and ang_var carried across.
Inputs are resident in HBM before timing starts.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, os.path.join(REPO, "mclics-backend_amd"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

L = 256
TEMPS = (1.2, 1.13, 1.0, 0.6)
BYTES_PER_SITE_UPDATE = 24  # one FP32 director (3x4 B) read and written per sweep
HBM_PEAK_GBS = 8000.0
METRIC = "site-updates/s at 256³ LL; decorrelated sweeps/s learned-proposal vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """(threads, description) for the CPU legs: OMP_NUM_THREADS when it is
    set (a scheduler may set it to the process's CPU share), else every core in
    this process's affinity mask; with the host CPU model."""
    avail = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = env if env > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"cpu_model": model, "affinity_cores": avail, "online_cpus": os.cpu_count(),
                     "threads_from": "OMP_NUM_THREADS" if env > 0 else "sched_getaffinity"}


def cpu_baseline(n0, budget_s=15.0):
    """The oracle's OpenMP restatement of the reference CPU sweep
    (src/evolve.cpp:64-179, 8 sublattice passes) on the host cores, timed on a
    bounded number of 256^3 sweeps at T=1.2."""
    import _oracle as O
    threads, host = host_cores()
    p = O.params(L, L, L, A=1.0)
    n = n0.copy()
    pt = np.ones(L**3, np.int32)
    av = 0.5
    t0 = time.perf_counter()
    _, av = O.sweep_rng(p, n, pt, None, 1.2, av, seed=2, sweep=0, nthreads=threads)
    first = time.perf_counter() - t0
    reps = max(1, min(60, int(budget_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    for w in range(reps):
        _, av = O.sweep_rng(p, n, pt, None, 1.2, av, seed=2, sweep=1 + w, nthreads=threads)
    dt = time.perf_counter() - t0
    rate = reps * L**3 / dt
    # one sweep on a single core, for the per-core figure
    t0 = time.perf_counter()
    _, av = O.sweep_rng(p, n, pt, None, 1.2, av, seed=2, sweep=1000, nthreads=1)
    one = L**3 / (time.perf_counter() - t0)
    return {"value": rate, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": f"{reps} sweeps of the 256^3 LL lattice at T=1.2 (after 1 warm-up sweep), "
                      f"reference CPU order (8 sublattice passes), OpenMP static schedule, same Philox proposals",
            "value_1core": one, **host}


def chain_stats(lat, sweep_us, budget_s, target=0.2, min_sweeps=20000):
    """E and S of a running chain, measured until tau_int(S) is resolved to
    `target` relative error (and the chain holds min_sweeps sweeps) or the
    budget in seconds is spent.  A first 20,000-sweep series with every sweep
    measured sets the thinning k ~ tau/10 (1..200); tau in sweeps is then
    k tau_int of the thinned series.  Returns a dict of the chain's figures."""
    from mclics.decorrelation import tau_error, tau_int
    t0 = time.perf_counter()
    E0, S0 = lat.measure_series(20000)
    tau0, _ = tau_int(S0)
    every = int(min(200, max(1, tau0 // 10)))
    Es, Ss = ([E0], [S0]) if every == 1 else ([], [])
    # samples per call: about two seconds of sweeps (measurement ~16 us)
    per = max(500, int(2.0e6 / (every * sweep_us + 16.0)))
    last = t0
    while True:
        if Ss:
            S = np.concatenate(Ss)
            tau, M = tau_int(S)
            err = tau_error(tau, M, len(S))
            done = err <= target * tau and len(S) * every >= min_sweeps
            now = time.perf_counter()
            if done or now - t0 > budget_s:
                break
            if now - last > 30.0:
                log(f"bench:   chain at {len(S) * every} sweeps: tau_int(S) {tau * every:.1f} +- {err * every:.1f}")
                last = now
        E1, S1 = lat.measure_series(per, every=every)
        Es.append(E1)
        Ss.append(S1)
    E = np.concatenate(Es)
    tauE, _ = tau_int(E)
    tau_s = max(tau, tauE)  # samples; for the error of the means
    n = len(S)
    return {"every": every, "samples": n, "sweeps": n * every, "tau_int_S": tau * every, "tau_err": err * every,
            "rel_err": err / tau, "resolved": bool(err <= target * tau), "window": M * every,
            "E": float(E.mean()), "S": float(S.mean()),
            "E_sem": float(E.std(ddof=1) * np.sqrt(2.0 * tau_s / n)),
            "S_sem": float(S.std(ddof=1) * np.sqrt(2.0 * tau_s / n)), "seconds": time.perf_counter() - t0}


def learned_leg(budget_s=30.0, n_train=1_000_000, L5=128, temps=(1.10, 1.12, 1.14), epochs=40, schedule=None,
                target=0.2, reflect=True):
    """Config C5: 128^3 LL near the transition.  The network is trained on
    n_train neighbourhoods drawn from plain-sampler lattices at the three
    temperatures, then plain and learned chains start from the same
    equilibrated state; tau_int(S) by automatic windowing gives decorrelated
    sweeps/s = sweeps/s / (2 tau_int).  Each chain runs until tau_int(S) is
    known to `target` relative error or `budget_s` seconds are spent; the gain carries a 95 %
    interval from the two relative errors, and E and S of the two chains are
    compared at 3 sigma (standard errors of the means from tau_int)."""
    import mclics
    nt = L5**3
    out = {"workload": f"LL bulk {L5}^3 periodic, A=1, T in {list(temps)}", "n_train": n_train,
           "moves": "vMF draw, then reflection through the learned mean" if reflect else "vMF draw only",
           "budget_s_per_chain": budget_s, "tau_rel_err_target": target, "per_T": []}
    starts = {}
    with mclics.Network(seed=5) as net:
        per_T = n_train // len(temps)
        t0 = time.perf_counter()
        for ti, T in enumerate(temps):
            lat = mclics.Lattice(mclics.Config(L5, L5, L5, T=T, seed=50 + ti), mclics.random_ic(L5, L5, L5, seed=50 + ti))
            lat.sweeps(3000, graph=True)
            snaps = 4
            for k in range(snaps):
                lat.sweeps(50, graph=True)
                net.add_samples(lat, per_T // snaps + (n_train - per_T * len(temps) if ti == 0 and k == 0 else 0),
                                seed=1000 * ti + k)
            lat.sweeps(20000, graph=True)  # further equilibration of the chains' common start
            starts[T] = lat.directors()
            lat.close()
        # (epochs, learning rate) stages; Adam's moments and step carry across stages
        losses = np.concatenate([net.train(epochs=e, batch=2048, lr=lr)
                                 for e, lr in (schedule or [(epochs, 2e-3)])])
        out["train_seconds"] = time.perf_counter() - t0
        out["loss_per_epoch"] = [float(v) for v in losses]
        gains = []
        for ti, T in enumerate(temps):
            row = {"T": T}
            for method in ("plain", "learned"):
                with mclics.Lattice(mclics.Config(L5, L5, L5, T=T, seed=90 + ti), starts[T]) as lat:
                    if method == "learned":
                        lat.set_learned(net, reflect=reflect)
                        lat.sweeps(200, graph=True)
                        ms, _ = lat.time_sweeps(1000, graph=True)
                        sps = 1000 / (ms / 1e3)
                    else:
                        lat.autotune()  # the plain chain at its fastest kernel shape for this size
                        lat.sweeps(200, graph=True)
                        ms, _ = lat.time_sweeps(1000, graph=True)
                        sps = 1000 / (ms / 1e3)
                    st = chain_stats(lat, 1e6 / sps, budget_s, target)
                    acc, _ = lat.acceptance()
                st.update({"sweeps_per_s": sps, "decorrelated_sweeps_per_s": sps / (2 * st["tau_int_S"]),
                           "acceptance": acc / nt})
                row[method] = st
                log(f"bench: learned leg T={T} {method}: {sps:.1f} sweeps/s tau={st['tau_int_S']:.1f} "
                    f"+- {st['tau_err']:.1f} over {st['sweeps']} sweeps (every {st['every']}) acc={acc / nt:.3f}")
            p, q = row["plain"], row["learned"]
            row["gain"] = q["decorrelated_sweeps_per_s"] / p["decorrelated_sweeps_per_s"]
            rel = float(np.hypot(p["rel_err"], q["rel_err"]))
            row["gain_95"] = [row["gain"] * float(np.exp(-1.96 * rel)), row["gain"] * float(np.exp(1.96 * rel))]
            row["resolved"] = p["resolved"] and q["resolved"]
            for what in ("E", "S"):
                d = abs(p[what] - q[what])
                sig = float(np.hypot(p[what + "_sem"], q[what + "_sem"]))
                row[what + "_agree_3sigma"] = bool(d <= 3 * sig)
                row[what + "_diff_over_sigma"] = d / sig if sig > 0 else float("inf")
            flops = 11520.0 * nt * q["sweeps_per_s"]
            q["mlp_tflops"] = flops / 1e12
            q["mfma_util_bf16_dense"] = flops / 2.5e15
            gains.append(row["gain"])
            out["per_T"].append(row)
        out["gain_geomean"] = float(np.exp(np.mean(np.log(gains))))
    return out


def batch_leg(n_sweeps=400):
    """Config C4: 64 independent periodic LL replicas (16 each of 32^3, 48^3,
    64^3 and 96^3; 20,643,840 sites), T_r uniform in [0.9, 1.3], IC seeds
    100..163, swept together (one batched launch per sweep, replayed as a
    graph) at the shape the batch's autotuner picks."""
    import mclics
    sizes = [32] * 16 + [48] * 16 + [64] * 16 + [96] * 16
    rng = np.random.default_rng(0)
    cfgs = [mclics.Config(s, s, s, T=float(np.float32(rng.uniform(0.9, 1.3))), seed=100 + r, stream=r)
            for r, s in enumerate(sizes)]
    ics = [mclics.random_ic(s, s, s, seed=100 + r) for r, s in enumerate(sizes)]
    tot = sum(s**3 for s in sizes)
    with mclics.Batch(cfgs, ics) as b:
        # the batch's own autotuner: segment-kernel grids and slab tile shapes,
        # timed on scratch copies (the chains are untouched)
        tuned_grid, _, tuned_us = b.autotune()
        b.sweeps(20)
        b.sync()
        ms = b.time_sweeps(n_sweeps)
        kernel = {mclics.KERNEL_SLAB: "k_batch_slab", mclics.KERNEL_X4: "k_batch_x4"}.get(b.kernel(), "?")
        E, S = b.measure()
        n_ms = 50  # measured sweeps: a sweep, then E and S of every replica
        b.measure_series(n_ms)  # allocates the series buffer
        b.sync()
        t0 = time.perf_counter()
        b.measure_series(n_ms)
        measured_us = (time.perf_counter() - t0) / n_ms * 1e6
    rate = n_sweeps * tot / (ms / 1e3)
    return {"workload": "64 LL replicas periodic: 16 each of 32^3, 48^3, 64^3, 96^3; T_r ~ U[0.9, 1.3]",
            "sites": tot, "sweeps": n_sweeps, "us_per_sweep": ms * 1e3 / n_sweeps, "kernel": kernel,
            "site_updates_per_s": rate, "roofline_frac_24B": rate * BYTES_PER_SITE_UPDATE / (HBM_PEAK_GBS * 1e9),
            "measured_sweep_us": measured_us, "measured_sweep_note": "eager sweep then E, S of every replica",
            "autotuned_workgroups": tuned_grid, "autotune_best_us": tuned_us,
            "S_mean": float(np.mean(S)), "E_mean": float(np.mean(E))}


def float_loop(xi, xf, dx):
    """The reference's float loop, for (x = xi; (int)1e6*sign*(x - xf) >= 0;
    x += dx) in FP32 with sign = (int)(-dx/|dx|) (src/evolve_thermal.cu:69,79)."""
    xs, x = [], np.float32(xi)
    sign = int(-dx / abs(dx))
    while np.float32(1000000 * sign) * (x - np.float32(xf)) >= 0:
        xs.append(float(x))
        x = np.float32(x + np.float32(dx))
    return xs


def c1_run(run_cpu, mct=500, mcs=200):
    """Config C1, the reference's own CPU-runnable case: 30^3 LL periodic,
    random start (seed 1), thermal 1.2 -> 0.0 by -0.02 (61 rows of the float
    loop), MCT = 500 sweeps then MCS = 200 measured sweeps (E and S after
    each) per temperature.  The whole run on the GPU; with run_cpu, the same
    run with the oracle's OpenMP sweep (8 sublattice passes, FP64 E and Q) on
    the host cores."""
    import mclics
    L1 = 30
    temps = float_loop(1.2, 0.0, -0.02)
    n0 = mclics.random_ic(L1, L1, L1, seed=1)
    out = {"workload": "LL 30^3 periodic, thermal 1.2 -> 0.0 by -0.02, MCT=500, MCS=200", "rows": len(temps),
           "site_updates": len(temps) * (mct + mcs) * L1**3}
    with mclics.Lattice(mclics.Config(L1, L1, L1, T=temps[0], seed=1), n0) as lat:
        lat.sweeps(2)
        lat.sync()
        t0 = time.perf_counter()
        rows = []
        for T in temps:
            lat.set_T(T)
            lat.sweeps(mct, graph=True)
            E, S = lat.measure_series(mcs)
            rows.append((T, float(S.mean()), float(E.mean())))
        out["gpu_seconds"] = time.perf_counter() - t0
        out["kernel"] = {mclics.KERNEL_SLAB: "k_sweep_slab", mclics.KERNEL_X4: "k_sweep_x4"}.get(lat.kernel(),
                                                                                                  "k_sweep_pass")
    out["gpu_site_updates_per_s"] = out["site_updates"] / out["gpu_seconds"]
    out["last_row"] = {"T": rows[-1][0], "S": rows[-1][1], "E": rows[-1][2]}
    if run_cpu:
        # the first rows of the same run, as many as fit in about budget_s
        import _oracle as O
        budget_s = 12.0
        threads, _ = host_cores()
        p = O.params(L1, L1, L1, A=1.0)
        n = n0.copy()
        pt = np.ones(L1**3, np.int32)
        av, sweep, done = 0.5, 0, 0
        t0 = time.perf_counter()
        for T in temps:
            for w in range(mct + mcs):
                _, av = O.sweep_rng(p, n, pt, None, np.float32(T), av, seed=1, sweep=sweep, nthreads=threads)
                sweep += 1
                if w >= mct:
                    O.energy_mean(p, n, pt, None, gpu=1, nthreads=threads)
                    O.eigvals(O.qtensor(n, pt))
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
        out["cpu_rows_run"] = done
        out["cpu_cores"] = threads
        out["cpu_site_updates_per_s"] = done * (mct + mcs) * L1**3 / dt
        out["cpu_seconds_full_run"] = out["site_updates"] / out["cpu_site_updates_per_s"]
    return out


def slab_geometry_leg(n_sweeps=400, p0=0.0):
    """Config C3: 128x128x64 GHRL (k11, k22, k33 = 1.0, 0.7, 1.3; p0 = 0, or
    p0 = 16 for the cholesteric variant) between homeotropic GHRL anchoring
    faces (W = 5), field along z (elecA = 1, elecE = 0.5), T = 0.5, periodic
    x/y and free z, homogeneous start along x.  Sweeps are timed after
    autotuning the fast kernels; 25 B per site-update (the 24 B of the
    directors plus the 1 B point type of the slab layout)."""
    import mclics
    from mclics.ic import homogeneous_ic
    nx, ny, nz = 128, 128, 64
    nt = nx * ny * nz
    k11, k22, k33 = 1.0, 0.7, 1.3
    S = abs(9.0 / (k11 - 3 * k22 - k33))  # src/io.cpp:282-288
    sigma = -S * k22 * (2 * np.pi / p0) if p0 else 0.0
    ghrl = (S / 9 * (k11 - k33), S / 9 * (2 * k11 - 3 * k22 + k33), S * (k22 - k11), S / 9 * (k11 - 3 * k22 - k33),
            sigma)
    pt = np.ones(nt, np.int32)
    ns = np.zeros(3 * nt, np.float32)
    pt[: nx * ny] = 2
    pt[nx * ny * (nz - 1):] = 3
    ns[2: 3 * nx * ny: 3] = -1.0
    ns[3 * nx * ny * (nz - 1) + 2::3] = 1.0
    surf = [mclics.Surface("homeotropic_ghrl", 5.0, (0.0, 0.0, 1.0))] * 2
    cfg = mclics.Config(nx, ny, nz, bound=(mclics.PERIODIC, mclics.PERIODIC, mclics.FREE), potential=mclics.GHRL,
                        ghrl=[float(np.float32(v)) for v in ghrl], elec=(0.0, 0.0, 1.0), elecA=1.0, elecE=0.5,
                        surfaces=surf, T=0.5, seed=7)
    with mclics.Lattice(cfg, homogeneous_ic(nx, ny, nz, 90.0, 0.0), pt=pt, ns_aos=ns) as lat:
        lat.autotune()
        lat.sweeps(20, graph=True)
        ms, _ = lat.time_sweeps(n_sweeps, graph=True)
        kernel = {mclics.KERNEL_SLAB: "k_sweep_slab", mclics.KERNEL_X4: "k_sweep_x4"}.get(lat.kernel(), "k_sweep_pass")
        E, S_, _ = lat.measure()
    rate = n_sweeps * nt / (ms / 1e3)
    return {"workload": f"GHRL 128x128x64 p0={p0:g}, periodic x/y, free z, homeotropic_ghrl faces W=5, field z "
                        "elecA=1 elecE=0.5, T=0.5", "sites": nt, "sweeps": n_sweeps, "us_per_sweep": ms * 1e3 / n_sweeps,
            "kernel": kernel, "site_updates_per_s": rate,
            "roofline_frac_25B": rate * 25 / (HBM_PEAK_GBS * 1e9), "E": E, "S": S_}


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` at 256^3 from the committed byte-counter
    summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by the copy
    calibration), or None."""
    path = os.path.join(REPO, "profiles", f"{kernel}_ll256_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("hbm_bytes_per_launch")
    except Exception:
        return None


def copy_probe(nbytes=201326592, reps=20):
    """Device-to-device copies of a known size, the calibration run for the
    FETCH_SIZE / WRITE_SIZE counters (each copy reads and writes nbytes)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def stream_ceiling(nbytes=3 * 4 * 256**3, reps=100):
    """Ping-pong device copies of one 256^3 plane set (A -> B, B -> A, ...):
    every copy reads what the previous one wrote, as consecutive sweeps do,
    so neither side is left in the 256 MB last-level cache from a repeat of
    the same input.  Gives the read+write rate the HBM sustains for the
    sweep's bytes without any update work."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 0, C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(b, 0, C.c_size_t(nbytes)) == 0
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    d2d = 3  # hipMemcpyDeviceToDevice

    def pingpong(n):
        for i in range(n):
            src, dst = (a, b) if i % 2 == 0 else (b, a)
            assert hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), d2d, None) == 0

    pingpong(20)
    assert hip.hipEventRecord(e0, None) == 0
    pingpong(reps)
    assert hip.hipEventRecord(e1, None) == 0
    assert hip.hipEventSynchronize(e1) == 0
    ms = C.c_float()
    assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    us = ms.value * 1e3 / reps
    tbs = 2.0 * nbytes / (us * 1e-6) / 1e12
    return {"method": "hipMemcpyAsync device-to-device ping-pong of one 256^3 plane set (3 x 67.1 MB)",
            "copies": reps, "us_per_copy": us, "read_plus_write_TB_s": tbs,
            "frac_of_8TBs": tbs / (HBM_PEAK_GBS / 1e3),
            "sweep_floor_us": us, "note": "a sweep moves the same bytes, so this copy time is the floor a "
                                          "sweep could reach with every update hidden under memory traffic"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-replica leg")
    ap.add_argument("--no-slab-geometry", action="store_true", help="skip the GHRL slab-geometry leg")
    ap.add_argument("--no-c1", action="store_true", help="skip the 30^3 thermal run")
    ap.add_argument("--learned-epochs", type=int, default=40, help="training epochs of the learned-proposal leg")
    ap.add_argument("--learned-budget", type=float, default=30.0,
                    help="seconds per chain of the learned-proposal leg (each chain stops earlier once "
                         "tau_int(S) is known to --learned-target)")
    ap.add_argument("--learned-target", type=float, default=0.2,
                    help="relative error of tau_int(S) at which a chain of the learned-proposal leg stops")
    ap.add_argument("--learned-no-reflect", action="store_true",
                    help="learned sweeps make the vMF move only (no reflection through the learned mean)")
    ap.add_argument("--kernel", choices=("auto", "slab", "x4", "generic"), default="auto",
                    help="sweep kernel of the headline leg (auto: slab where it fits)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the device-copy calibration for the HBM byte counters")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU z-slab decomposition is a design, not a code path)", file=sys.stderr)
        sys.exit(2)
    if args.steps < len(TEMPS):
        print("bench.py: --steps must be at least the number of temperatures (4)", file=sys.stderr)
        sys.exit(2)

    import mclics
    if mclics.device_count() < 1:
        print("bench.py: no gfx950 device visible", file=sys.stderr)
        sys.exit(1)
    nt = L**3
    log("bench: generating 256^3 random directors (seed 2)")
    n0 = mclics.random_ic(L, L, L, seed=2)
    cfg = mclics.Config(L, L, L, potential=mclics.LL, A=1.0, T=TEMPS[0], seed=2)
    per_t = [args.steps // len(TEMPS) + (1 if i < args.steps % len(TEMPS) else 0) for i in range(len(TEMPS))]
    obs = []
    with mclics.Lattice(cfg, n0) as lat:
        lat.set_kernel({"auto": mclics.KERNEL_AUTO, "slab": mclics.KERNEL_SLAB, "x4": mclics.KERNEL_X4,
                        "generic": mclics.KERNEL_GENERIC}[args.kernel])
        log(f"bench: warmup {args.warmup} sweeps")
        if args.warmup:
            lat.sweeps(args.warmup)
        lat.sync()
        # Each temperature's sweeps run back to back between one HIP event
        # pair on the lattice's stream: the mean launch duration is the
        # events' time over the launches (kernels plus the gaps between
        # them).  An event pair around every launch would add its own few
        # microseconds to each and stretch the wall clock.
        region_ms, passes = 0.0, 0
        t0 = time.perf_counter()
        for T, k in zip(TEMPS, per_t):
            lat.set_T(T)
            _, tot_ms, npass = lat.time_passes(k, 0)
            region_ms += tot_ms
            passes += npass
        lat.sync()
        wall = time.perf_counter() - t0
        for T in TEMPS[-1:]:
            E, S, _ = lat.measure()
            obs.append({"T": T, "E": E, "S": S})
        acc, sweeps = lat.acceptance()
        # Measured sweeps (a sweep then E and S, as in a run's MCS phase),
        # outside the timed region: their cost beside a plain sweep.
        n_ms = 200
        lat.sync()
        t1 = time.perf_counter()
        lat.sweeps(n_ms)
        lat.sync()
        t2 = time.perf_counter()
        lat.measure_series(n_ms)
        t3 = time.perf_counter()
        measured_step = {"sweeps": n_ms, "plain_sweep_us": (t2 - t1) / n_ms * 1e6,
                         "measured_sweep_us": (t3 - t2) / n_ms * 1e6,
                         "measurement_us": ((t3 - t2) - (t2 - t1)) / n_ms * 1e6,
                         "note": "eager launches; measurement = E (FP64 sum of site energies) and the "
                                 "FP64 Q tensor with its largest eigenvalue S"}
        kind = lat.kernel()
        kname = {mclics.KERNEL_SLAB: "k_sweep_slab (slab kernel, both colours, one launch per sweep)",
                 mclics.KERNEL_X4: "k_sweep_x4 (segment kernel, one colour pass)"}.get(
            kind, "k_sweep_pass (site kernel, one colour pass)")
        tile = lat.slab_tile() if kind == mclics.KERNEL_SLAB else None
        log(f"bench: {args.steps} sweeps in {wall:.3f} s, mean pass {region_ms / passes * 1e3:.2f} us")
    value = nt * args.steps / wall
    avg_pass_s = region_ms / passes / 1e3
    # a launch updates one colour (two-pass kernels) or the whole lattice (slab)
    launches_per_sweep = passes / args.steps
    bytes_per_launch = BYTES_PER_SITE_UPDATE * nt / launches_per_sweep
    achieved = bytes_per_launch / avg_pass_s / 1e9
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox-seeded directors uniform on the sphere (seed 2)",
        "config": {"workload": "Lebwohl-Lasher bulk 256x256x256 periodic, A=1, checkerboard sweeps over "
                               "T in {1.2, 1.13, 1.0, 0.6}", "lattice": [L, L, L], "temperatures": list(TEMPS),
                   "sweeps_per_T": per_t, "order": "checkerboard (2 passes/sweep)", "seed": 2},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(kname.split()[0]),
                     "kernel": kname, "bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": avg_pass_s * 1e6, "launches_per_sweep": launches_per_sweep,
                     "launches_timed": passes, "timing": "HIP event pair per temperature block "
                                                        "of back-to-back launches",
                     "event_ms": region_ms,
                     "slab_tile": tile},
        "stream_ceiling": stream_ceiling(),
        "observables": obs,
        "measured_sweeps": measured_step,
        "last_acceptance": acc / nt,
    }
    if not args.no_batch:
        log("bench: batched-replica leg (64 replicas, 20.6M sites)")
        result["batched_replicas"] = batch_leg()
    if not args.no_slab_geometry:
        log("bench: slab-geometry leg (GHRL 128x128x64, anchoring faces, field)")
        result["slab_geometry"] = slab_geometry_leg()
        log("bench: slab-geometry leg, cholesteric variant (p0 = 16)")
        result["slab_geometry_cholesteric"] = slab_geometry_leg(p0=16.0)
    if not args.no_learned:
        log("bench: learned-proposal leg (128^3, T near T_NI)")
        result["learned_proposal"] = learned_leg(args.learned_budget, epochs=args.learned_epochs,
                                                 target=args.learned_target, reflect=not args.learned_no_reflect)
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, OpenMP)")
        result["cpu_baseline"] = cpu_baseline(n0)
    if not args.no_c1:
        log("bench: C1 run (30^3 thermal, 61 rows) on the GPU" + ("" if args.no_cpu else " and the host cores"))
        result["c1_thermal_run"] = c1_run(not args.no_cpu)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
