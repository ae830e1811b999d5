"""Benchmark: trimem energy + gradient evaluations per second on one MI355X.

Workload (the single-mesh precision configuration): icosphere level 7
(163842 vertices, 327680 faces), radius 1, vertices jittered by
N(0, (0.01 * mean edge)^2) from seed 20261203; kappa_b=30, kappa_a=1e6,
kappa_v=1e6, kappa_c=0, kappa_t=1e5, kappa_r=1e3, area fraction 1.0, volume
fraction 0.8, Edge tether (r=2, lc0=1.25 l, lc1=0.75 l), cell-list repulsion
with exclusion level 2 and rc = rlist = 1.5 l.  One step = one energy + gradient
evaluation (properties pass + gradient pass) on positions already resident in
HBM, in the mixed precision path by default.

Beside the headline, the same JSON line carries blocks for device-resident
HMC steps (`hmc`), the surrogate force network (`surrogate`), replica groups
(`replicas`), the serial and parallel flip sweeps (`flips`), 32 L7 meshes in
one group (`combined`) and the CPU oracle's rates (`cpu_baseline`, with its
serial flip sweep under `flips`); each has a flag that switches it off.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC_FULL = "energy+grad evals/s & HBM GB/s; surrogate vs plain HMC effective samples/s"
HBM_PEAK_GBPS = 8000.0


def workload(level):
    from trimem_amd import workloads
    return workloads.single(level)


PROFILE_JSON = os.path.join(ROOT, "profiles", "pmc_energy_L7_mixed.json")
COPY_BYTES = 64 << 20


class Hip:
    """hipEvent timing on the library's stream."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                          ctypes.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def copy_calibration(self, stream, nbytes=COPY_BYTES):
        """One device-to-device copy of known size (the runtime's copy
        kernel), so that HBM counters of the same run can be scaled."""
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(a), nbytes) == 0
        assert self.h.hipMalloc(ctypes.byref(b), nbytes) == 0
        assert self.h.hipMemcpyAsync(b, a, nbytes, 3, stream) == 0    # hipMemcpyDeviceToDevice
        assert self.h.hipStreamSynchronize(stream) == 0
        self.h.hipFree(a)
        self.h.hipFree(b)

    def event(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        t = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(t), a, b) == 0
        return float(t.value)

    def destroy(self, e):
        self.h.hipEventDestroy(e)

    def marked_pass1_us(self, L, run, k):
        """Mean duration (us) of the pass-1 kernel over k calls of run(),
        each a properties pass whose pass-1 kernel stamps a pair of events at
        its own start and end (tm_mark_pass1)."""
        ev = [(self.event(), self.event()) for _ in range(k)]
        for a, b in ev:
            assert L.tm_mark_pass1(a, b) == 0
            run()
        assert L.tm_sync() == 0
        d = [self.elapsed_ms(a, b) * 1e3 for a, b in ev]
        for a, b in ev:
            self.destroy(a)
            self.destroy(b)
        return float(np.mean(d))


def cpu_baseline(v, f, params, budget_s):
    """The FP64 OpenMP oracle (the reference's vertex-centric structure) on
    the same inputs: one eval = energy() + gradient() as the reference's
    callers issue them (2 properties passes + 1 gradient pass)."""
    from oracle import oracle as O
    om = O.Mesh(v, f)
    oe = O.EnergyManager(om, O.Params(**params))
    for _ in range(2):
        oe.energy(om)
        oe.gradient(om)
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < 3 or (time.perf_counter() < t_end and len(times) < 50):
        t0 = time.perf_counter()
        oe.energy(om)
        oe.gradient(om)
        times.append(time.perf_counter() - t0)
    return 1.0 / float(np.median(times)), len(times), O.num_threads()


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_flip_baseline(v, f, params, ratio=0.1, seed=7):
    """The oracle's serial flip sweep (flips.cpp:22-80, one thread) on the
    same mesh: one sweep of ratio * edges proposals, injected uniformly at
    random with their uniforms; attempted flips per second."""
    from oracle import oracle as O
    om = O.Mesh(v, f)
    oe = O.EnergyManager(om, O.Params(**params))
    ne = len(f) * 3 // 2
    n = int(ne * ratio)
    rng = np.random.default_rng(seed)
    prop = rng.integers(0, ne, n).astype(np.int32)
    uni = rng.random(n)
    t0 = time.perf_counter()
    acc, used = oe.flip_serial_injected(om, prop, uni)
    dt = time.perf_counter() - t0
    return {"attempted_flips_per_s": n / dt, "ms_per_sweep": 1e3 * dt, "accepted": int(acc), "proposals": n,
            "cores": 1}


def core_params(core, params):
    from trimem_amd import workloads
    return workloads.core_params(core, params)


def hmc_rates(v, f, params, precision, dev_steps, host_steps):
    """HMC steps/s of the same chain two ways: the device-resident step
    (DeviceMeshHMC, device draws) and the drop-in path (MeshHMC over
    EnergyEvaluators over trimem_amd.core, numpy draws, positions through
    the host view on every force call).  Both: dt 5e-6 (acceptance ~0.95 at
    L7; at 1e-5 most trajectories diverge on the stiff tethers), 10 leapfrog steps,
    mass 1, T 1, neighbour list rebuilt after every step."""
    import contextlib
    import io
    from trimem_amd import core
    from trimem_amd.mc import evaluators as EV
    from trimem_amd.mc import hmc as H
    from trimem_amd.mc.mesh import Mesh

    opts = {"time_step": 5e-6, "num_integration_steps": 10, "info_step": 0}

    def device_chain(pair_cache):
        mesh = Mesh(v, f)
        est = core.EnergyManager(mesh.trimesh, core_params(core, params))
        est.precision = precision
        tun = est.tuning
        tun["pair_cache"] = pair_cache
        est.tuning = tun
        with contextlib.redirect_stdout(io.StringIO()):
            dev = H.DeviceMeshHMC(mesh, est, counter=H.get_step_counters(), options=opts, rng="device", seed=1)
        dev.step()
        est.update_repulsion(mesh.trimesh)
        a0 = dev.result()["n_accepted"]
        t0 = time.perf_counter()
        for _ in range(dev_steps):
            dev.step()
            est.update_repulsion(mesh.trimesh)
        r = dev.result()
        return time.perf_counter() - t0, r["n_accepted"] - a0

    t_dev, acc = device_chain(1)
    t_full, _ = device_chain(0)

    mesh2 = Mesh(v, f)
    est2 = core.EnergyManager(mesh2.trimesh, core_params(core, params))
    est2.precision = precision
    counter = H.get_step_counters()
    with contextlib.redirect_stdout(io.StringIO()):
        funcs = EV.EnergyEvaluators(mesh2, est2, None, {"info_step": 0, "output_step": 0, "refresh_step": 1})
        host = H.MeshHMC(mesh2, funcs.fun, funcs.grad, counter=counter, options=opts)
    host.x = mesh2.x.copy()
    np.random.seed(1)
    host.step()
    funcs.callback(host.x, counter)
    t0 = time.perf_counter()
    for _ in range(host_steps):
        host.step()
        funcs.callback(host.x, counter)
    t_host = time.perf_counter() - t0
    return {"device_steps_per_s": dev_steps / t_dev, "device_steps_per_s_full_query": dev_steps / t_full,
            # energy+gradient evaluations at positions that move between
            # evaluations, the leapfrog kernels and the per-step list rebuild
            # included: the conservative counterpart of the headline rate
            "moving_evals_per_s": 11 * dev_steps / t_dev,
            "dropin_steps_per_s": host_steps / t_host,
            "device_over_dropin": (dev_steps / t_dev) / (host_steps / t_host),
            "device_acceptance": acc / dev_steps,
            "device_steps": dev_steps, "dropin_steps": host_steps,
            "trajectory": "dt 5e-6, 10 leapfrog steps (11 force evaluations), mass 1, T 1, "
                          "neighbour list rebuilt every step"}


def replica_workload(n_replicas=64, seed=20261204):
    from trimem_amd import workloads
    return workloads.replicas(n_replicas, seed)


def replica_rates(precision, evals, hmc_steps, hip, stream):
    """Batched-replica rates: one grouped evaluation of all replicas (every
    kernel one launch over all meshes) against the same meshes evaluated one
    by one, and grouped HMC steps (dt 1e-5, 10 leapfrog steps, device draws,
    replica exchange every 10 steps) with and without hipGraph replay."""
    from trimem_amd import capi, core
    L = capi.lib()
    work, ladder = replica_workload()
    meshes, ests = [], []
    for v, f, params in work:
        m = core.TriMesh(v, f)
        e = core.EnergyManager(m, core_params(core, params))
        e.precision = precision
        meshes.append(m)
        ests.append(e)
    R = core.Replicas(meshes, ests)
    tune = R.autotune()
    nv = np.array([m.n_vertices() for m in meshes])
    nf = np.array([m.n_faces() for m in meshes])
    eval_bytes = float(np.sum((24.0 * nv + 12.0 * nf + 16.0 * nv) * 2 + 24.0 * nv))

    def timed(fn, k):
        fn()
        capi.check(L.tm_sync())
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        for _ in range(k):
            fn()
        hip.record(e1, stream)
        return hip.elapsed_ms(e0, e1) / k

    ms_group = timed(lambda: R.eval(), evals)
    handles = [(ctypes.c_void_p(capi_handle(m)), ctypes.c_void_p(capi_handle(e))) for m, e in zip(meshes, ests)]
    both = capi.TM_EVAL_PROPS | capi.TM_EVAL_GRAD

    def one_by_one():
        for mh, eh in handles:
            L.tm_energy_eval_device(eh, mh, both)
    ms_seq = timed(one_by_one, max(2, evals // 4))

    R.hmc_setup(mass=1.0, time_step=1e-5, temperature=1.0, seeds=list(range(1, len(meshes) + 1)),
                num_integration_steps=10)
    R.set_ladder(list(ladder), seed=7)
    ms_nograph = timed(lambda: R.hmc_run(10, swap_every=10, use_graph=False), max(1, hmc_steps // 10)) / 10
    ms_graph = timed(lambda: R.hmc_run(10, swap_every=10, use_graph=True), max(1, hmc_steps // 10)) / 10
    st = R.stats()
    return {"replicas": len(meshes), "vertices": int(nv.sum()), "levels": sorted(set(int(x) for x in
                                                                                   np.round(np.log(
                                                                                       (nv - 2) / 10) / np.log(4)))),
            "group_evals_per_s": 1e3 / ms_group, "replica_evals_per_s": len(meshes) * 1e3 / ms_group,
            "hbm_gbps": eval_bytes / (ms_group * 1e-3) / 1e9, "algorithmic_bytes_per_group_eval": eval_bytes,
            "one_by_one_group_evals_per_s": 1e3 / ms_seq, "grouped_over_one_by_one": ms_seq / ms_group,
            "hmc_group_steps_per_s_graph": 1e3 / ms_graph, "hmc_group_steps_per_s_direct": 1e3 / ms_nograph,
            "hmc_acceptance": float(st["accepted"].sum() / max(st["steps"].sum(), 1)),
            "swap_acceptance": float(st["swaps_accepted"].sum() / max(st["swaps_tried"].sum(), 1)),
            "tuning": {k: tune[k] for k in ("rep_threads", "rep_batch", "pass2_threads", "rep_fused", "pair_cache")},
            "workload": "64 icosphere meshes, levels 3-6 from seed, vfrac U[0.6,0.95], kappa_b log-uniform "
                        "[10,100] ladder; HMC dt 1e-5, 10 leapfrog steps, exchange every 10 steps"}


def flip_rates(v, f, params, precision, ratio=0.1, sweeps=3):
    """Attempted edge flips per second at L7 (flip_ratio 0.1 of the edges per
    sweep): the serial sweep (one workgroup walks the proposals in order, as
    flips.cpp:22-80 does) and the parallel sweep (claim-based conflict-free
    rounds, flips.cpp:82-166)."""
    from trimem_amd import core
    out = {"ratio": ratio}
    for kind, fn in (("serial", core.flip), ("parallel", core.pflip)):
        mesh = core.TriMesh(v, f)
        est = core.EnergyManager(mesh, core_params(core, params))
        est.precision = precision
        n = int(mesh.n_edges() * ratio)
        fn(mesh, est, ratio, seed=1)
        t0 = time.perf_counter()
        acc = sum(fn(mesh, est, ratio, seed=2 + k) for k in range(sweeps))
        dt = time.perf_counter() - t0
        out[kind] = {"attempted_flips_per_s": sweeps * n / dt, "ms_per_sweep": 1e3 * dt / sweeps,
                     "accepted_per_sweep": acc / sweeps, "proposals_per_sweep": n}
    out["edges"] = int(mesh.n_edges())
    return out


def combined_rates(precision, hip, stream, n_replicas=32, evals=20, seed=20261205, chain_steps=20, swap_every=5,
                   time_step=5e-6, flip_ratio=0.2, guarded_steps=16):
    """The combined configuration on one GPU: 32 icosphere L7 meshes
    (163842 vertices each, jittered 1% from their own seeds), volume
    fractions U[0.55, 0.9], kappa_b log-uniform on [20, 80] as a tempering
    ladder.

    * grouped energy+gradient evaluations of all 32 (every kernel one launch
      over all meshes) and their algorithmic HBM rate;
    * the Monte Carlo chain: each step is one grouped device HMC step of
      every replica (10 leapfrog steps, device draws) followed by a parallel
      flip sweep over 20% of the edges of every replica; a kappa_b exchange
      round every swap_every steps;
    * the surrogate's plain-gradient fallback on the first replica: guarded
      HMC steps driven by an untrained network, which the policy (window 5,
      acceptance 0.5) replaces by the exact gradient."""
    from trimem_amd import capi, core, workloads
    L = capi.lib()
    work, ladder = workloads.combined(n_replicas, seed)
    ms, es = [], []
    for v, f, params in work:
        m = core.TriMesh(v, f)
        e = core.EnergyManager(m, core_params(core, params))
        e.precision = precision
        ms.append(m)
        es.append(e)
    R = core.Replicas(ms, es)
    tune = R.autotune()
    nv = np.array([m.n_vertices() for m in ms], dtype=np.float64)
    nf = np.array([m.n_faces() for m in ms], dtype=np.float64)
    per_vertex = 104.0 if precision == "fp64" else 52.0    # measurement model, B = per_vertex N + 24 F
    eval_bytes = float(np.sum(per_vertex * nv + 24.0 * nf))
    R.eval()
    capi.check(L.tm_sync())
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(evals):
        R.eval()
    hip.record(e1, stream)
    ms_group = hip.elapsed_ms(e0, e1) / evals
    # the two long kernels of the group, each timed live: pass 2 alone (the
    # gradient over pass 1's outputs of the last evaluation), back to back,
    # and the pass-1 kernel between events stamped at its start and end;
    # the roofline line is the longer one
    wide = precision == "fp64"
    p1_bytes = float(np.sum((40.0 if wide else 20.0) * nv + 12.0 * nf))
    p2_bytes = float(np.sum((64.0 if wide else 32.0) * nv + 12.0 * nf))
    hip.record(e0, stream)
    for _ in range(evals):
        R.eval(properties=False, gradient=True)
    hip.record(e1, stream)
    capi.check(L.tm_sync())
    us_p2 = hip.elapsed_ms(e0, e1) / evals * 1e3
    us_p1 = hip.marked_pass1_us(L, lambda: R.eval(properties=True, gradient=False), evals)
    R.eval()
    capi.check(L.tm_sync())
    sfx = "_grouped_fp64" if wide else "_grouped"
    lines = [{"kernel": "k_pass1_apart" + sfx, "bound": "hbm", "achieved": p1_bytes / us_p1 / 1e3,
              "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": p1_bytes / us_p1 / 1e3 / HBM_PEAK_GBPS,
              "algorithmic_bytes": p1_bytes, "duration_us": us_p1},
             {"kernel": "k_pass2" + sfx, "bound": "hbm", "achieved": p2_bytes / us_p2 / 1e3,
              "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": p2_bytes / us_p2 / 1e3 / HBM_PEAK_GBPS,
              "algorithmic_bytes": p2_bytes, "duration_us": us_p2}]
    gp = os.path.join(ROOT, "profiles", "pmc_energy_32xL7_grouped.json")
    if os.path.exists(gp) and n_replicas == 32:
        ks = json.load(open(gp))["kernels"]
        for r in lines:
            kk = ks.get(r["kernel"], {})
            r["traffic"] = kk.get("hbm_bytes_per_dispatch")
            r["rocprof_median_us"] = kk.get("avg_us")
    roof = max(lines, key=lambda r: r["duration_us"])
    other_roof = min(lines, key=lambda r: r["duration_us"])
    out = {"replicas": n_replicas, "vertices": int(nv.sum()), "group_evals_per_s": 1e3 / ms_group,
           "ms_per_group_eval": ms_group, "replica_evals_per_s": n_replicas * 1e3 / ms_group,
           "algorithmic_bytes_per_group_eval": eval_bytes, "hbm_gbps": eval_bytes / (ms_group * 1e-3) / 1e9,
           "hbm_frac": eval_bytes / (ms_group * 1e-3) / 1e9 / HBM_PEAK_GBPS,
           "roofline": roof, "roofline_other_pass": other_roof,
           "tuning": {k: tune[k] for k in ("rep_threads", "rep_batch", "pass2_threads", "pair_cache")},
           "workload": "32 icosphere L7 meshes jittered 1% (own seeds), vfrac U[0.55,0.9], kappa_b "
                       "log-uniform [20,80], cell-list repulsion excl 2 rc=1.5l; one grouped energy+gradient "
                       "evaluation of all"}

    if chain_steps <= 0:
        return out
    # the chain: grouped HMC + flip sweeps + exchanges
    R.hmc_setup(mass=1.0, time_step=time_step, temperature=1.0, seeds=list(range(1, n_replicas + 1)),
                num_integration_steps=10)
    R.set_ladder(list(ladder), seed=11)
    R.hmc_run(1)                  # warm-up (graph-free, every kernel shape loaded)
    capi.check(L.tm_sync())
    st0 = R.stats()
    t_hmc = t_flip = 0.0
    flips_acc, flips_tried = 0, 0
    for k in range(chain_steps):
        t0 = time.perf_counter()
        R.hmc_run(1, swap_every=1 if (k + 1) % swap_every == 0 else 0)
        capi.check(L.tm_sync())
        t1 = time.perf_counter()
        # one parallel sweep of every replica, round by round over all of them
        flips_acc += sum(R.pflip(flip_ratio, [1000 * k + i for i in range(n_replicas)]))
        flips_tried += sum(int(m.n_edges() * flip_ratio) for m in ms)
        capi.check(L.tm_sync())
        t_hmc += t1 - t0
        t_flip += time.perf_counter() - t1
    st = R.stats()
    acc = float((st["accepted"] - st0["accepted"]).sum() / max((st["steps"] - st0["steps"]).sum(), 1))
    out["chain"] = {"steps": chain_steps, "steps_per_s": chain_steps / (t_hmc + t_flip),
                    "hmc_group_steps_per_s": chain_steps / t_hmc,
                    "flip_attempts_per_s": flips_tried / t_flip,
                    "hmc_acceptance": acc,
                    "swap_acceptance": float(st["swaps_accepted"].sum() / max(st["swaps_tried"].sum(), 1)),
                    "flip_acceptance": flips_acc / max(flips_tried, 1),
                    "setup": f"HMC dt {time_step}, 10 leapfrog steps, T 1, device draws; parallel flip sweep of "
                             f"{flip_ratio:.0%} of the edges of every replica (all replicas in one sweep) after "
                             f"every HMC step; one kappa_b exchange round every {swap_every} steps"}

    # the surrogate's fallback on replica 0
    sg = core.Surrogate(ms[0], es[0], seed=3)
    sg.set_policy(window=5, min_acceptance=0.5, probe_every=5)
    d = core.DeviceHMC(ms[0], es[0])
    drivers, accepted = [], []
    t0 = time.perf_counter()
    for k in range(guarded_steps):
        drivers.append(d.step(mass=1.0, time_step=time_step, num_integration_steps=10, temperature=1.0, seed=77,
                              step=k, surrogate=sg, guarded=True))
        accepted.append(bool(d.result()["accepted"]))
    dt_g = time.perf_counter() - t0
    pol = sg.policy_state()
    out["surrogate_fallback"] = {
        "steps": guarded_steps, "drivers": drivers, "accepted": accepted, "steps_per_s": guarded_steps / dt_g,
        "surrogate_steps": pol["surrogate_steps"], "exact_steps": pol["exact_steps"],
        "falling_back": pol["falling_back"], "switches": pol["switches"],
        "exact_acceptance": float(np.mean([a for dr, a in zip(drivers, accepted) if dr == "exact"] or [0.0])),
        "setup": "replica 0, untrained network, policy window 5 / min acceptance 0.5 / probe every 5 steps"}
    return out


def ess(x):
    """Effective sample size of a scalar chain (Geyer's initial monotone
    sequence estimator of the integrated autocorrelation time)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    x = x - x.mean()
    var = float(x @ x) / n
    if var == 0.0:
        return 1.0    # a chain that never moved holds one sample
    if n < 4:
        return float(n)
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / (n * var)
    tau, prev = -1.0, np.inf
    for k in range(0, n - 1, 2):
        g = acf[k] + acf[k + 1]
        if g <= 0:
            break
        g = min(g, prev)
        tau += 2.0 * g
        prev = g
    return n / max(tau, 1.0)


def surrogate_rates(precision, level=6, n_train=120, epochs=15, chain=300, time_step=1e-5, L=10, lr=1e-3):
    """Surrogate-driven against plain HMC on one mesh: training from
    snapshots of the plain chain (GPU forward/backward kernels, Adam), the
    held-out force error, and for equal-length chains from the same state
    steps/s, acceptance and the effective sample size of the energy."""
    from trimem_amd import core
    v, f, lbar, params = workload(level)
    mesh = core.TriMesh(v, f)
    est = core.EnergyManager(mesh, core_params(core, params))
    est.precision = precision
    d = core.DeviceHMC(mesh, est)
    kw = dict(mass=1.0, time_step=time_step, num_integration_steps=L, temperature=1.0)
    sg = core.Surrogate(mesh, est, seed=1)
    t0 = time.perf_counter()
    for s_ in range(n_train):
        d.step(seed=3, step=s_, **kw)
        sg.add_snapshot()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = sg.train(epochs=epochs, lr=lr, seed=1)
    t_train = time.perf_counter() - t0
    # held-out state: a few more plain steps
    for s_ in range(n_train, n_train + 5):
        d.step(seed=3, step=s_, **kw)
    F = sg.forces()
    G = est.gradient(mesh)
    force_err = float(np.linalg.norm(F + G) / np.linalg.norm(G))
    start = mesh.points().copy()

    def run(use_sg, seed):
        mesh.points()[:] = start
        est.energy(mesh)    # host positions are the start state again
        trace, acc = [], 0
        t0 = time.perf_counter()
        for k in range(chain):
            d.step(seed=seed, step=k, surrogate=sg if use_sg else None, **kw)
            r = d.result()
            acc += int(r["accepted"])
            trace.append(r["energy1"] if r["accepted"] else r["energy0"])
        t = time.perf_counter() - t0
        e = ess(trace)
        return {"steps_per_s": chain / t, "acceptance": acc / chain, "ess": e, "ess_per_s": e / t}

    plain = run(False, 11)
    surr = run(True, 11)
    return {"workload": f"icosphere L{level} ({len(v)} vertices), benchmark parameters, HMC dt {time_step}, "
                        f"{L} leapfrog steps, T 1; {n_train} training snapshots from the plain chain, "
                        f"{epochs} epochs of Adam (lr {lr})",
            "train_seconds": t_train, "data_seconds": t_data, "final_train_loss": loss,
            "heldout_force_rel_err": force_err, "plain": plain, "surrogate": surr,
            "ess_per_s_ratio": surr["ess_per_s"] / plain["ess_per_s"] if plain["ess_per_s"] > 0 else None}


def capi_handle(obj):
    """The C handle behind a trimem_amd.core TriMesh / EnergyManager."""
    return obj._handle()


def _csv_rows(path):
    import csv
    return list(csv.DictReader(open(path)))


def _find(d, suffix):
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith(suffix):
                return os.path.join(root, f)
    raise FileNotFoundError(f"{suffix} under {d}")


def _short(name):
    if "copyBuffer" in name:
        return "copy"
    for k in ("k_rep_cached", "k_rep_record", "k_rep", "k_pass1_apart", "k_pass1", "k_pass2",
              "k_finalize", "k_generic_props", "k_generic_grad"):
        if k in name:
            return k + ("_grouped" if "_grouped" in name else "") + ("_fp64" if "<double" in name else "")
    return None


# the launches of pass 1 with the pair cache replaying (the default)
PASS1_KERNELS = ("k_pass1_apart", "k_rep_cached", "k_finalize")
SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS",
               "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES", "SQ_ACTIVE_INST_VALU")
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4    # 1024 SIMDs, one wave64 VALU op per 4 cycles at 2.4 GHz


L7_N, L7_F = 163842, 327680
# (title, the kernels a roofline line may be about with their algorithmic
# bytes per launch and how they are made up, the launches of pass 1); the
# summary's line is about the longer of them
SINGLE_PROFILE = ("energy + gradient passes, icosphere L7, mixed precision",
                  (("k_pass1_apart", 20.0 * L7_N + 12.0 * L7_F, f"20 N + 12 F, N = {L7_N}, F = {L7_F}"),
                   ("k_pass2", 32.0 * L7_N + 12.0 * L7_F, f"32 N + 12 F, N = {L7_N}, F = {L7_F}")),
                  PASS1_KERNELS)
GROUP_PROFILE = ("energy + gradient passes, 32 icosphere L7 meshes in one group, mixed precision",
                 (("k_pass1_apart_grouped", 32 * (20.0 * L7_N + 12.0 * L7_F),
                   f"32 meshes x (20 N + 12 F), N = {L7_N}, F = {L7_F}"),
                  ("k_pass2_grouped", 32 * (32.0 * L7_N + 12.0 * L7_F),
                   f"32 meshes x (32 N + 12 F), N = {L7_N}, F = {L7_F}")),
                 tuple(k + "_grouped" for k in PASS1_KERNELS))


def summarize(trace_dir, fetch_dir, write_dir, out_json, out_md, command, sq_dir=None, what=None):
    """Summaries of rocprofv3 runs of this script: average kernel durations
    (--kernel-trace --stats) and HBM KiB per dispatch (--pmc FETCH_SIZE, then
    WRITE_SIZE), scaled by what the counters read for the calibration copy
    of COPY_BYTES; with sq_dir, the VALU instructions per dispatch
    (--pmc SQ_COUNTERS) and the time their issue takes."""
    # per kernel the median over its dispatches (a recording launch among
    # replays, or the autotuner's trial shapes, do not move it)
    durs = {}
    for r in _csv_rows(_find(trace_dir, "kernel_trace.csv")):
        k = _short(r["Kernel_Name"])
        if k:
            durs.setdefault(k, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    stats = {k: {"calls": len(d), "avg_us": float(np.median(d)), "mean_us": float(np.mean(d))}
             for k, d in durs.items()}
    pmc, calib = {}, {}
    for d, cname in ((fetch_dir, "FETCH_SIZE"), (write_dir, "WRITE_SIZE")):
        for r in _csv_rows(_find(d, "counter_collection.csv")):
            k = _short(r.get("Kernel_Name", r.get("Kernel-Name", "")))
            if not k or r["Counter_Name"] != cname:
                continue
            if k == "copy":
                # the runtime issues small copies of its own; calibrate on
                # the COPY_BYTES dispatch, the one with the largest grid
                calib.setdefault(cname, []).append((int(r["Grid_Size"]), float(r["Counter_Value"])))
            else:
                pmc.setdefault(k, {}).setdefault(cname, []).append(float(r["Counter_Value"]))
    copy_kib = COPY_BYTES / 1024.0
    f_fetch = max(calib["FETCH_SIZE"])[1] / copy_kib
    f_write = max(calib["WRITE_SIZE"])[1] / copy_kib
    kernels = {}
    for k, c in pmc.items():
        fk = float(np.median(c["FETCH_SIZE"]))
        wk = float(np.median(c["WRITE_SIZE"]))
        kernels[k] = {"fetch_kib": fk, "write_kib": wk,
                      "hbm_bytes_per_dispatch": (fk / f_fetch + wk / f_write) * 1024.0,
                      "avg_us": stats.get(k, {}).get("avg_us")}
    if sq_dir:
        sq = {}
        for r in _csv_rows(_find(sq_dir, "counter_collection.csv")):
            k = _short(r.get("Kernel_Name", r.get("Kernel-Name", "")))
            if k and k != "copy":
                sq.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
        for k, c in sq.items():
            if k in kernels and "SQ_INSTS_VALU" in c:
                valu = float(np.median(c["SQ_INSTS_VALU"]))
                kernels[k]["valu_insts"] = valu
                kernels[k]["valu_issue_us"] = valu / VALU_ISSUE_PER_S * 1e6
                kernels[k]["waves"] = float(np.median(c.get("SQ_WAVES", [0])))
    out = {"command": command, "copy_bytes": COPY_BYTES, "copy_factor_fetch": f_fetch,
           "copy_factor_write": f_write, "kernels": kernels, "durations": stats}
    json.dump(out, open(out_json, "w"), indent=1)
    title, roofs, pass1 = what or SINGLE_PROFILE
    lines = ["# rocprofv3 summary: " + title, "",
             "Commands (run from the repository root on one MI355X):", "", "```"] + command + ["```", "",
             f"Copy calibration: {COPY_BYTES} B device-to-device copy reads FETCH_SIZE x {f_fetch:.4f} "
             f"and WRITE_SIZE x {f_write:.4f} of its size; kernel traffic below is divided by these factors.", "",
             "Durations and counters are medians over each kernel's dispatches.", "",
             "| kernel | calls | duration (us) | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM bytes / dispatch |",
             "|---|---|---|---|---|---|"]
    for k in sorted(stats):
        kk = kernels.get(k, {})
        lines.append(f"| {k} | {stats[k]['calls']} | {stats[k]['avg_us']:.2f} | "
                     f"{kk.get('fetch_kib', float('nan')):.1f} | {kk.get('write_kib', float('nan')):.1f} | "
                     f"{kk.get('hbm_bytes_per_dispatch', float('nan')):.0f} |")
    # mixed precision per vertex: positions 12, a and c 8, gradient 12 B; 12 B per face
    have = sorted((r for r in roofs if r[0] in stats), key=lambda r: -stats[r[0]]["avg_us"])
    for i, (rk, rbytes, rtext) in enumerate(have):
        us = stats[rk]["avg_us"]
        which = "the longest kernel" if i == 0 else "the other pass"
        lines += ["", f"Roofline of {rk} ({which}): {rbytes:.0f} algorithmic bytes per launch "
                      f"({rtext}) / {us:.2f} us = {rbytes / us / 1e3:.0f} GB/s, "
                      f"{rbytes / us / 1e3 / HBM_PEAK_GBPS:.3f} of the 8 TB/s peak."]
    p1 = [k for k in pass1 if k in stats]
    if p1:
        lines += ["", f"Pass 1 ({' + '.join(p1)}): {sum(stats[k]['avg_us'] for k in p1):.2f} us, "
                      f"{sum(kernels.get(k, {}).get('hbm_bytes_per_dispatch', 0.0) for k in p1):.0f} HBM bytes."]
    if sq_dir:
        lines += ["", "VALU issue (--pmc " + " ".join(SQ_COUNTERS) + "): wave64 VALU instructions per dispatch "
                  "and the time their issue takes at one instruction per 4 cycles on each of 1024 SIMDs "
                  "at 2.4 GHz.", "",
                  "| kernel | waves | VALU insts | VALU insts / wave | VALU issue (us) | issue / duration |",
                  "|---|---|---|---|---|---|"]
        for k in sorted(kernels):
            kk = kernels[k]
            if "valu_insts" in kk and stats.get(k):
                lines.append(f"| {k} | {kk['waves']:.0f} | {kk['valu_insts']:.0f} | "
                             f"{kk['valu_insts'] / max(kk['waves'], 1):.0f} | {kk['valu_issue_us']:.2f} | "
                             f"{kk['valu_issue_us'] / stats[k]['avg_us']:.2f} |")
    open(out_md, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


SG_COUNTERS = ("SQ_INSTS_VALU_MFMA_MOPS_BF16", "MfmaUtil")
BF16_MFMA_PEAK_TFLOPS = 2500.0    # MI355X dense bf16


def _sg_name(name):
    """k_sg_* kernel name with its template arguments, or None."""
    i = name.find("k_sg_")
    if i < 0:
        return None
    j = name.find("(", i)
    return name[i:j if j > 0 else None]


def summarize_surrogate(trace_dir, pmc_dir, out_md, command):
    """Summary of the surrogate network's kernels: durations (--kernel-trace
    --stats) and, per dispatch, the bf16 MFMA work (--pmc
    SQ_INSTS_VALU_MFMA_MOPS_BF16, 512 FLOPs each) and the MFMA busy share
    of the SIMDs (the derived MfmaUtil)."""
    stats = {}
    for r in _csv_rows(_find(trace_dir, "kernel_stats.csv")):
        k = _sg_name(r["Name"])
        if k:
            stats[k] = {"calls": int(r["Calls"]), "avg_us": float(r["AverageNs"]) / 1e3,
                        "total_ms": float(r["TotalDurationNs"]) / 1e6}
    pmc = {}
    for r in _csv_rows(_find(pmc_dir, "counter_collection.csv")):
        k = _sg_name(r.get("Kernel_Name", r.get("Kernel-Name", "")))
        if k:
            pmc.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    lines = ["# rocprofv3 summary: surrogate network kernels, icosphere L6 (40962 vertices)", "",
             "Commands (run from the repository root on one MI355X):", "", "```"] + command + ["```", "",
             "bf16 MFMA work per dispatch from SQ_INSTS_VALU_MFMA_MOPS_BF16 (units of 512 FLOPs), rate over the "
             f"kernel's average duration, against the {BF16_MFMA_PEAK_TFLOPS:.0f} TFLOP/s dense bf16 peak; "
             "MfmaUtil is rocprofv3's derived MFMA-busy share of all SIMD cycles while the kernel runs.", "",
             "| kernel | calls | avg (us) | total (ms) | MFMA GFLOP / dispatch | TFLOP/s | of peak | MfmaUtil (%) |",
             "|---|---|---|---|---|---|---|---|"]
    total = sum(v["total_ms"] for v in stats.values())
    for k in sorted(stats, key=lambda k: -stats[k]["total_ms"]):
        st, c = stats[k], pmc.get(k, {})
        fl = float(np.mean(c["SQ_INSTS_VALU_MFMA_MOPS_BF16"])) * 512 if "SQ_INSTS_VALU_MFMA_MOPS_BF16" in c else 0.0
        util = float(np.mean(c["MfmaUtil"])) if "MfmaUtil" in c else float("nan")
        tf = fl / (st["avg_us"] * 1e-6) / 1e12 if fl else 0.0
        lines.append(f"| {k} | {st['calls']} | {st['avg_us']:.2f} | {st['total_ms']:.2f} | {fl / 1e9:.4f} | "
                     f"{tf:.2f} | {tf / BF16_MFMA_PEAK_TFLOPS:.4f} | {util:.2f} |")
    lines += ["", f"All surrogate kernels: {total:.2f} ms over the run."]
    open(out_md, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--surrogate", type=int, default=1,
                    help="1: include the surrogate-vs-plain HMC block (L6), 0: skip")
    ap.add_argument("--replica-evals", type=int, default=40,
                    help="grouped evaluations of the 64-replica set timed for the replicas block (0: skip)")
    ap.add_argument("--sg-train", type=int, default=120, help="surrogate block: training snapshots")
    ap.add_argument("--sg-epochs", type=int, default=15, help="surrogate block: training epochs")
    ap.add_argument("--sg-chain", type=int, default=300, help="surrogate block: HMC steps per chain")
    ap.add_argument("--flips", type=int, default=1, help="attempted flips/s of serial and parallel sweeps")
    ap.add_argument("--combined", type=int, default=1, help="grouped evaluations of 32 L7 meshes")
    ap.add_argument("--combined-chain", type=int, default=20,
                    help="combined block: Monte Carlo steps of the chain (0: evaluations only)")
    ap.add_argument("--summarize-group", nargs=4, metavar="DIR",
                    help="TRACE_DIR FETCH_DIR WRITE_DIR SQ_DIR of the combined block's evaluations: write the "
                         "grouped passes' summary to profiles/ and exit")
    ap.add_argument("--hmc-steps", type=int, default=40,
                    help="device-resident HMC steps timed for the hmc block (0: skip)")
    ap.add_argument("--summarize-surrogate", nargs=2, metavar="DIR",
                    help="TRACE_DIR PMC_DIR: write the surrogate kernels' summary to profiles/ and exit")
    ap.add_argument("--summarize", nargs="+", metavar="DIR",
                    help="TRACE_DIR FETCH_DIR WRITE_DIR [SQ_DIR]: write profiles/ summaries from the "
                         "rocprofv3 output folders and exit")
    args = ap.parse_args()
    # the OpenMP oracle (loaded later) pinned one thread per core, close
    # together, as the measurement model asks for the CPU baseline
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if args.summarize_group:
        grp = (" --steps 10 --warmup 2 --no-cpu-baseline --hmc-steps 0 --replica-evals 0 --surrogate 0 --flips 0"
               " --combined 1 --combined-chain 0")
        d = args.summarize_group
        cmd = ["rocprofv3 --kernel-trace --stats --output-format csv -d " + d[0] + " -o run -- python3 bench.py" + grp,
               "rocprofv3 --pmc FETCH_SIZE --output-format csv -d " + d[1] + " -o run -- python3 bench.py" + grp,
               "rocprofv3 --pmc WRITE_SIZE --output-format csv -d " + d[2] + " -o run -- python3 bench.py" + grp,
               "rocprofv3 --pmc " + " ".join(SQ_COUNTERS) + " --output-format csv -d " + d[3] +
               " -o run -- python3 bench.py" + grp,
               "python3 bench.py --summarize-group " + " ".join(d)]
        summarize(d[0], d[1], d[2], os.path.join(ROOT, "profiles", "pmc_energy_32xL7_grouped.json"),
                  os.path.join(ROOT, "profiles", "energy_passes_32xL7_grouped.md"), cmd, d[3], GROUP_PROFILE)
        return
    if args.summarize_surrogate:
        sg = (" --steps 10 --warmup 2 --no-cpu-baseline --hmc-steps 0 --replica-evals 0 --flips 0 --combined 0"
              " --surrogate 1 --sg-train 40 --sg-epochs 5 --sg-chain 20")
        cmd = ["rocprofv3 --kernel-trace --stats --output-format csv -d " + args.summarize_surrogate[0] +
               " -o run -- python3 bench.py" + sg,
               "rocprofv3 --pmc " + " ".join(SG_COUNTERS) + " --output-format csv -d " +
               args.summarize_surrogate[1] + " -o run -- python3 bench.py" + sg,
               "python3 bench.py --summarize-surrogate " + " ".join(args.summarize_surrogate)]
        summarize_surrogate(args.summarize_surrogate[0], args.summarize_surrogate[1],
                            os.path.join(ROOT, "profiles", "surrogate_kernels_L6.md"), cmd)
        return
    if args.summarize:
        only = " --no-cpu-baseline --hmc-steps 0 --replica-evals 0 --surrogate 0 --flips 0 --combined 0"
        cmd = ["rocprofv3 --kernel-trace --stats --output-format csv -d " + args.summarize[0] +
               " -o run -- python3 bench.py --steps 200 --warmup 10" + only,
               "rocprofv3 --pmc FETCH_SIZE --output-format csv -d " + args.summarize[1] +
               " -o run -- python3 bench.py --steps 20 --warmup 2" + only,
               "rocprofv3 --pmc WRITE_SIZE --output-format csv -d " + args.summarize[2] +
               " -o run -- python3 bench.py --steps 20 --warmup 2" + only]
        if len(args.summarize) > 3:
            cmd.append("rocprofv3 --pmc " + " ".join(SQ_COUNTERS) + " --output-format csv -d " +
                       args.summarize[3] + " -o run -- python3 bench.py --steps 20 --warmup 2" + only)
        cmd.append("python3 bench.py --summarize " + " ".join(args.summarize))
        summarize(args.summarize[0], args.summarize[1], args.summarize[2], PROFILE_JSON,
                  os.path.join(ROOT, "profiles", "energy_passes_L7_mixed.md"), cmd,
                  args.summarize[3] if len(args.summarize) > 3 else None)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: every box here has a single MI355X; "
              "multi-GPU sharding is a design note, not a launch mode", file=sys.stderr)
        sys.exit(2)

    from trimem_amd import capi
    L = capi.lib()
    v, f, lbar, params = workload(args.level)
    n, nf = len(v), len(f)

    mesh = ctypes.c_void_p()
    vv = np.ascontiguousarray(v)
    ff = np.ascontiguousarray(f, dtype=np.int32)
    capi.check(L.tm_mesh_create(vv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                ff.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), nf, ctypes.byref(mesh)))
    cp = capi.tm_params()
    for k, _ in capi.tm_params._fields_:
        if k != "reserved":
            setattr(cp, k, params[k])
    em = ctypes.c_void_p()
    capi.check(L.tm_energy_create(mesh, ctypes.byref(cp), ctypes.byref(em)))
    capi.check(L.tm_energy_set_precision(em, capi.TM_PREC_MIXED if args.precision == "mixed" else capi.TM_PREC_FP64))
    capi.check(L.tm_mesh_upload_points(mesh))
    stream = ctypes.c_void_p()
    capi.check(L.tm_stream(ctypes.byref(stream)))
    hip = Hip()
    hip.copy_calibration(stream)
    both = capi.TM_EVAL_PROPS | capi.TM_EVAL_GRAD

    for _ in range(args.warmup):
        capi.check(L.tm_energy_eval_device(em, mesh, both))
    capi.check(L.tm_sync())

    # timed region: K evaluations end to end
    e0, e1 = hip.event(), hip.event()
    capi.check(L.tm_sync())
    t0 = time.perf_counter()
    hip.record(e0, stream)
    for _ in range(args.steps):
        L.tm_energy_eval_device(em, mesh, both)
    hip.record(e1, stream)
    capi.check(L.tm_sync())
    wall = time.perf_counter() - t0
    ms_total = hip.elapsed_ms(e0, e1)
    evals_per_s = args.steps / (ms_total / 1e3)

    # per-launch durations, events on the library stream around a run of
    # back-to-back launches of one pass (a pair of events around every single
    # launch would add the event's own cost to each): the gradient pass is
    # one kernel (k_pass2; the mesh is closed, so no boundary walker follows
    # it) and reads what the last properties pass left; the properties pass
    # is the launch sequence k_pass1_apart + k_rep_cached + k_finalize
    ks = min(args.steps, 500)

    def per_launch_ms(flags, k):
        hip.record(e0, stream)
        for _ in range(k):
            L.tm_energy_eval_device(em, mesh, flags)
        hip.record(e1, stream)
        capi.check(L.tm_sync())
        return hip.elapsed_ms(e0, e1) / k

    p1 = np.array([per_launch_ms(capi.TM_EVAL_PROPS, ks)])
    p2 = np.array([per_launch_ms(capi.TM_EVAL_GRAD, ks)])
    # the pass-1 kernel alone (k_pass1_apart; the first of the properties
    # pass's launches), between events stamped at its start and end
    us_k1 = hip.marked_pass1_us(L, lambda: L.tm_energy_eval_device(em, mesh, capi.TM_EVAL_PROPS), ks)
    capi.check(L.tm_energy_eval_device(em, mesh, both))
    capi.check(L.tm_sync())

    def timed_evals(k):
        for _ in range(args.warmup):
            L.tm_energy_eval_device(em, mesh, both)
        hip.record(e0, stream)
        for _ in range(k):
            L.tm_energy_eval_device(em, mesh, both)
        hip.record(e1, stream)
        capi.check(L.tm_sync())
        return k / (hip.elapsed_ms(e0, e1) / 1e3)

    # the same evaluations with the pair cache off (the full cell-list query
    # every time), for comparison; the headline keeps the default
    tun = capi.tm_tuning()
    capi.check(L.tm_energy_get_tuning(em, ctypes.byref(tun)))
    tun.pair_cache = 0
    capi.check(L.tm_energy_set_tuning(em, ctypes.byref(tun)))
    full_query_evals_per_s = timed_evals(args.steps)
    tun.pair_cache = 1
    capi.check(L.tm_energy_set_tuning(em, ctypes.byref(tun)))
    # the other precision path on the same mesh
    other = "fp64" if args.precision == "mixed" else "mixed"
    capi.check(L.tm_energy_set_precision(em, capi.TM_PREC_FP64 if other == "fp64" else capi.TM_PREC_MIXED))
    other_evals_per_s = timed_evals(args.steps)
    capi.check(L.tm_energy_set_precision(em, capi.TM_PREC_MIXED if args.precision == "mixed" else capi.TM_PREC_FP64))

    # algorithmic bytes per launch, the measurement model's per-unit figures:
    # pass 1 reads positions and the triangle table (12 B/face) and writes
    # a_i, c_i; pass 2 reads the same plus a_i, c_i and writes the gradient.
    # Positions, a_i/c_i and the gradient count 24/16/24 B per vertex in the
    # FP64 path and 12/8/12 B in the mixed path (FP32 bulk), whatever width
    # the kernels move them in.
    wp, wq, wg = (24.0, 16.0, 24.0) if args.precision == "fp64" else (12.0, 8.0, 12.0)
    b_pass1 = wp * n + 12.0 * nf + wq * n
    b_pass2 = wp * n + 12.0 * nf + wq * n + wg * n
    t1, t2 = float(np.mean(p1)) * 1e-3, float(np.mean(p2)) * 1e-3
    eval_bytes = b_pass1 + b_pass2
    # the roofline line is the longer of the two long kernels: the pass-1
    # kernel (pass 1's bytes) and the pass-2 kernel (pass 2's bytes)
    sfx = "_fp64" if args.precision == "fp64" else ""
    prof = {}
    if os.path.exists(PROFILE_JSON) and args.level == 7:
        prof = json.load(open(PROFILE_JSON))["kernels"]

    def roof_line(kernel, nbytes, us):
        r = {"kernel": kernel, "bound": "hbm", "achieved": nbytes / us / 1e3, "peak": HBM_PEAK_GBPS,
             "unit": "GB/s", "frac": nbytes / us / 1e3 / HBM_PEAK_GBPS, "traffic": None,
             "algorithmic_bytes": nbytes, "duration_us": us, "rocprof_avg_us": None, "valu_issue_frac": None}
        kk = prof.get(kernel)
        if kk is not None:
            r["traffic"] = kk["hbm_bytes_per_dispatch"]
            r["rocprof_avg_us"] = kk.get("avg_us")
            if "valu_issue_us" in kk:
                # the time the kernel's VALU instructions take to issue / its duration
                r["valu_issue_frac"] = kk["valu_issue_us"] / us
        return r

    roofs = [roof_line("k_pass1_apart" + sfx, b_pass1, us_k1), roof_line("k_pass2" + sfx, b_pass2, t2 * 1e6)]
    roof = max(roofs, key=lambda r: r["duration_us"])
    other_roof = min(roofs, key=lambda r: r["duration_us"])
    out = {
        "metric": "energy+grad evals/s",
        "metric_full": METRIC_FULL,
        "value": evals_per_s,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"icosphere L{args.level} ({n} vertices, {nf} faces), jittered 1% of mean edge, "
                               "kappa_b=30 kappa_a=1e6 kappa_v=1e6 kappa_t=1e5 kappa_r=1e3, vfrac 0.8, "
                               "edge tether, cell-list repulsion excl 2 rc=1.5l; one step = energy+gradient",
                   "precision": args.precision, "vertices": n, "faces": nf},
        "hbm_gbps_per_eval": eval_bytes * evals_per_s / 1e9,
        "algorithmic_bytes_per_eval": eval_bytes,
        "kernels_ms": {"pass1_phase": t1 * 1e3, "k_pass1_apart" + sfx: us_k1 * 1e-3, "k_pass2" + sfx: t2 * 1e3},
        "roofline": roof,
        "roofline_other_pass": other_roof,
        "pass1_phase": {"kernels": list(PASS1_KERNELS), "duration_us": t1 * 1e6, "algorithmic_bytes": b_pass1,
                        "achieved_gbps": b_pass1 / t1 / 1e9, "frac": b_pass1 / t1 / 1e9 / HBM_PEAK_GBPS},
        "wall_s_timed": wall,
        "full_query_evals_per_s": full_query_evals_per_s,
        f"{other}_evals_per_s": other_evals_per_s,
    }
    if args.hmc_steps > 0:
        out["hmc"] = hmc_rates(v, f, params, args.precision, args.hmc_steps, max(2, args.hmc_steps // 20))
    if args.surrogate:
        out["surrogate"] = surrogate_rates(args.precision, n_train=args.sg_train, epochs=args.sg_epochs,
                                           chain=args.sg_chain)
    if args.replica_evals > 0:
        out["replicas"] = replica_rates(args.precision, args.replica_evals, args.hmc_steps, hip, stream)
    if args.flips:
        out["flips"] = flip_rates(v, f, params, args.precision)
    if args.combined:
        out["combined"] = combined_rates(args.precision, hip, stream, chain_steps=args.combined_chain)
    if not args.no_cpu_baseline:
        cv, ns, thr = cpu_baseline(v, f, params, args.cpu_seconds)
        out["cpu_baseline"] = {"value": cv, "unit": "evals/s", "cores": thr, "kind": "port",
                               "sample": f"{ns} FP64 energy()+gradient() evals of the same L{args.level} mesh, "
                                         f"median, OpenMP {thr} threads",
                               "cpu": cpu_model(), "host_cpus": os.cpu_count(),
                               "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND",
                                                                      "OMP_PLACES")}}
        out["speedup_vs_cpu"] = evals_per_s / cv
        if args.flips:
            out["cpu_baseline"]["flips"] = cpu_flip_baseline(v, f, params)
    print(json.dumps(out), flush=True)
    L.tm_energy_destroy(em)
    L.tm_mesh_destroy(mesh)


if __name__ == "__main__":
    main()
