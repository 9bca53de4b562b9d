#!/usr/bin/env python3
"""Throughput of the LJ cluster-nucleation step on one MI355X.

Default workload (ensemble512): 512 independent Nose-Hoover trajectories,
N log-uniform in 500..32768 (4.39 M atoms), rho* U(0.01,0.05), T* U(0.6,0.8),
cutoff 2.5 sigma, list radius 2.8 sigma, dt 0.005 tau, Stillinger analysis
every 100 steps with first-passage time to a 50-atom cluster.  A "step" is
one integrator step of every trajectory (rebuilds, forces, thermostat and the
scheduled cluster analyses included); data are synthetic (seeded vapour).

Timing: after W warmup steps (with a cluster analysis, so every kernel has
been loaded and the step graph captured), the autotuner (which recaptures the
graph) and one untimed analysis interval, exactly K steps are timed by the
host clock between two device synchronisations.  The analysis runs every
cluster_every steps; its cost (HIP events around each analysis of the untimed
interval) is added at that rate, K / cluster_every analyses, so that a short
window counts what a long run pays per step.

Reported:
  value            mixed-precision batched path, atom-steps/s: FP32 pair
                   arithmetic, each atom's pair forces accumulated in FP32 with
                   Kahan compensation (FCMD_MIXED_KAHAN), FP64 state,
                   integrator and reductions
  mixed_fp64_sums  the same run with each atom's FP32 pair forces summed in
                   FP64 instead (FCMD_MIXED): throughput, force error
  fp64_batched     the same run in FP64
  fp64_plain       the plain FP64 path: one trajectory at a time (sampled,
                   each on its own captured graph)
  speedup          value / fp64_plain
  accuracy         the mixed path against FP64: RMS relative force error on
                   the mixed run's own snapshots of the batch (FP64 forces
                   computed from the same positions), energy drift per tau
                   of both and their ratio, temperature of both, largest-
                   cluster series of both
  roofline         dominant kernel (the force kernel or the drift, whichever
                   is slower): its mean duration over K more steps, run
                   without the graph, from HIP events each launch records at
                   the kernel's start and end on the engine stream; traffic
                   from profiles/traffic_<kernel>_<workload>.json
  cpu_baseline     the reference Fortran (flang -O3) on the host cores: one
                   serial instance per core, the step loop timed by the
                   harness itself (the initial O(N^2) list build excluded and
                   reported apart)
  dtype            f32+f64: pair arithmetic in FP32, state, integrator and
                   reductions in FP64 (the fp64_* fields are FP64 throughout)
  tuning           launch tunables the autotuner picked during the warmup (or
                   those --tuning fixed)

Usage: python bench.py [--steps K] [--warmup W] [--workload NAME] [--gpus 1]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fortran-cluster-md_amd"))

import fcmd  # noqa: E402
from fcmd import workloads as W  # noqa: E402

HBM_PEAK = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_system(params, states, precision):
    s = fcmd.System(params, precision)
    W.setup(s, params, states)
    return s


def hamiltonian(s, params):
    """Per-trajectory conserved quantity ekin + epot + psi^2 Q/2 + 3NkT s."""
    out = []
    for t, p in enumerate(params):
        sc = s.scalars(t)
        nc = p.natom if not p.nh_mask else sum(p.num[a] for a in range(p.nsorts) if (p.nh_mask >> a) & 1)
        out.append(sc.ekin + sc.epot + sc.psi ** 2 * p.Q / 2 + 3 * nc * fcmd.CIKB * p.att * sc.s)
    return np.array(out)


def force_error(s_mixed, params, sample):
    """RMS relative force error of the mixed path against FP64 on the mixed
    run's own current state: for each sampled trajectory of the batch, FP64
    forces computed (neighbor(0) + AccAtom) from the same positions.
    Returns (pooled RMS, worst per-trajectory RMS)."""
    sub = [params[t] for t in sample]
    num = den = 0.0
    worst = 0.0
    with fcmd.System(sub, fcmd.FP64) as ref:
        got = []
        for k, t in enumerate(sample):
            st = s_mixed.download(t)
            got.append(st)
            n = params[t].natom
            ref.upload(dict(st, ax=np.zeros(n), ay=np.zeros(n), az=np.zeros(n),
                            api=_api(params[t])), traj=k)
            ref.neighbor(0, traj=k)
        ref.accatom()
        for k, t in enumerate(sample):
            a = ref.download(k)
            nk = sum(((got[k][q] - a[q]) ** 2).sum() for q in ("ax", "ay", "az"))
            dk = sum((a[q] ** 2).sum() for q in ("ax", "ay", "az"))
            num += nk
            den += dk
            worst = max(worst, float(np.sqrt(nk / dk)))
    return float(np.sqrt(num / den)), worst


def _api(p):
    return np.concatenate([np.full(p.num[s], s + 1, dtype=np.int32) for s in range(p.nsorts)])


def temperatures(s, params):
    """Per-trajectory temperature 2 ekin / (3 N k), ekin as the integrator
    defines it."""
    return np.array([2.0 * s.scalars(t).ekin / (3.0 * p.natom * fcmd.CIKB) for t, p in enumerate(params)])


def timed_run(s, K, every, use_graph=True, fpt_size=50):
    """The timed window: K steps between two synchronisations (every kernel
    already loaded, the graph captured), plus K / every analyses at the
    mean duration the events gave for the untimed interval before it.
    Returns (seconds, analysis ms, the untimed interval's series and fpt)."""
    series, fpt = s.run(every, cluster_every=every, fpt_size=fpt_size, use_graph=use_graph)
    an_ms, _ = s.last_run_analysis_ms()
    t0 = time.perf_counter()
    s.run(K, cluster_every=0, use_graph=use_graph)
    wall = time.perf_counter() - t0
    return wall + (K / every) * an_ms * 1e-3 if every else wall, an_ms, series, fpt


def measure(name, precisions, steps, warmup=10, cluster_every=None):
    """Throughput of one workload at each precision (same seeded start);
    energy drift per tau of the extended Hamiltonian; largest-cluster series."""
    params, states, meta = W.WORKLOADS[name]()
    natoms = int(sum(p.natom for p in params))
    every = meta.get("cluster_every", 0) if cluster_every is None else cluster_every
    out = {"atoms": natoms, "trajectories": len(params), "steps": steps}
    for prec in precisions:
        tag = "fp64" if prec == fcmd.FP64 else "mixed"
        with fcmd.System(params, prec) as s:
            W.setup(s, params, states)
            s.run(warmup, cluster_every=every, use_graph=True)
            s.autotune(rounds=3)
            h0 = hamiltonian(s, params)
            t_all, an_ms, series, _ = timed_run(s, steps, every)
            h1 = hamiltonian(s, params)
            tau = (steps + every) * params[0].dt / W.TAU
            out[tag] = {"atom_steps_per_s": natoms * steps / t_all, "ms_per_step": t_all / steps * 1e3,
                        "analysis_ms": an_ms,
                        "energy_drift_per_tau": float(np.median(np.abs(h1 - h0) / np.abs(h0) / tau)),
                        "temperature_K": float(temperatures(s, params).mean()),
                        "largest_cluster_series": series.max(axis=1).tolist() if series.size else []}
            if prec != fcmd.FP64 and fcmd.FP64 in precisions:
                out["rms_rel_force_err"], _ = force_error(s, params, [0])
        log(f"[bench] {name} {tag}: {out[tag]['atom_steps_per_s']:.4e} atom-steps/s")
    return out


# --------------------------------------------------------------------------
# CPU baseline: the reference's own plugin set (nose build: int-verlet-nh-mav,
# neighbor-pbc-mav-silent, for-a-lj-n-pbc-m, stoddard-pbc-n-m) compiled from
# /root/reference by oracle/Makefile, driven by oracle/ref_harness.f.
def host_cores():
    """Cores this process may run on, at most 16 (a one-GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(params, states, cores, budget_s=15.0):
    """One serial instance of the reference per core, each on one of the
    workload's trajectories (spread over the sizes), run side by side.  The
    harness times its actions itself: the step loop (integrator calls,
    including the reference's own O(N^2) list rebuilds) is what counts; the
    initial neighbor(0) is reported apart."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    builds = sorted(int(d[3:]) for d in os.listdir(os.path.join(po.REF_DIR, "bin"))
                    if d.startswith("mxa") and os.path.exists(po.harness_path("nose", int(d[3:]))))
    if not builds:
        return None
    order = np.argsort([p.natom for p in params])
    pick = [int(order[int(round(q))]) for q in np.linspace(0, len(order) - 1, cores)]
    procs = []
    tmp = tempfile.mkdtemp()
    for k, t in enumerate(pick):
        p, st = params[t], states[t]
        n = p.natom
        mxa = next((m for m in builds if m >= n), None)
        if mxa is None:
            continue
        # cost per step: O(N) update + O(N^2) list rebuild every ~12 steps
        cost = 2e-7 * n + 7.5e-9 * n * n / 2 / 12
        steps = int(np.clip(budget_s / cost, 10, 4000))
        case = dict(natom=n, nsorts=p.nsorts, api=st["api"], box=np.array(p.box[:]),
                    mas=np.array(p.mas[:p.nsorts]), sig=np.array(p.sig[:p.nsorts]),
                    eps=np.array(p.eps[:p.nsorts]), num=np.array(p.num[:p.nsorts]),
                    atrcut=p.atrcut, atrskin=p.atrskin, atskin=p.atskin, dt=p.dt, Q=p.Q,
                    att=p.att, nbtim=p.nbtim, btim=0, atnenum=p.atnenum, ethst_nh=1,
                    psi=W.nh_initial_psi(p, st), s=0.0, dmms=0.0, dmmax=0.0, dmmav=0.0,
                    dmmst=0, dmmi=0, **{q: st[q] for q in ("x", "y", "z", "vx", "vy", "vz",
                                                            "ax", "ay", "az")})
        w = float(np.float32(p.nbtim))
        case["dmmh"] = np.full(256, (p.atrskin - p.atrcut) / w)
        case["dmmsu"] = 256.0 * (p.atrskin - p.atrcut) / w
        fin = os.path.join(tmp, f"in{k}.bin")
        po._write_input(fin, case, [(1, 0, 0), (2, 0, 0), (3, 1, steps)])
        log_path = os.path.join(tmp, f"log{k}.txt")
        with open(log_path, "w") as lf:
            pr = subprocess.Popen([po.harness_path("nose", mxa), fin, os.path.join(tmp, f"out{k}.bin")],
                                  stdout=lf, stderr=subprocess.DEVNULL, preexec_fn=po._unlimited_stack)
        procs.append((pr, log_path, n, steps))
    rate = 0.0
    sizes, builds_s = [], []
    for pr, log_path, n, steps in procs:
        pr.wait()
        if pr.returncode != 0:
            continue
        times = {}
        for line in open(log_path):
            if line.startswith("HARNESS-TIME"):
                _, code, sec = line.split()
                times[int(code)] = float(sec)
        if 3 not in times or times[3] <= 0:
            continue
        rate += n * steps / times[3]
        sizes.append(n)
        builds_s.append(times.get(1, 0.0))
    return dict(value=rate, unit="atom-steps/s", cores=len(sizes), kind="reference",
                sample=f"{len(sizes)} of the trajectories (N={sorted(sizes)}), one serial instance "
                       f"per core on {len(sizes)} cores (this process's allowed cores, at most 16), "
                       f"~{budget_s:.0f} s of steps each, reference nose plugin set, flang -O3; "
                       "rate = sum over instances of N * steps / step-loop time (the harness's "
                       "own clock; list rebuilds inside the loop included, the initial "
                       "O(N^2) list build excluded)",
                initial_list_build_s_mean=float(np.mean(builds_s)) if builds_s else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="ensemble512", choices=sorted(W.WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="serial reference instances for the CPU baseline; 0: one per allowed core (at most 16)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step's kernels on the stream instead of replaying a graph")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (262k FP64, 2M FP64 vs mixed, mixture)")
    ap.add_argument("--tuning", default="",
                    help="fixed launch tunables force_u,nbr_cu,wl_grid instead of the autotuner "
                         "(for profiles whose passes must run the same kernels)")
    ap.add_argument("--nbtim", type=int, default=0,
                    help="override the deck's neighbour update time (nbtim, the adaptive skin's "
                         "target rebuild interval); 0 keeps the workload's")
    args = ap.parse_args()
    use_graph = not args.no_graph
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "trajectories shard across GPUs with no communication", file=sys.stderr)
        sys.exit(2)
    if fcmd.device_count() < 1:
        raise SystemExit("bench.py: no HIP device")

    K, Wm = args.steps, args.warmup
    log(f"[bench] building workload {args.workload}")
    params, states, meta = W.WORKLOADS[args.workload]()
    if args.nbtim:
        for p in params:
            p.nbtim = args.nbtim
    natoms = int(sum(p.natom for p in params))
    every = meta.get("cluster_every", 0)
    fpt_size = meta.get("fpt_size", 50)

    res = {}
    hdrift = {}
    sample = [int(t) for t in np.argsort([p.natom for p in params])[::max(1, len(params) // 16)]]
    for name, prec in (("mixed", fcmd.MIXED_KAHAN), ("mixed_fp64_sums", fcmd.MIXED), ("fp64", fcmd.FP64)):
        log(f"[bench] {name}: setup {len(params)} trajectories, {natoms} atoms")
        s = build_system(params, states, prec)
        # warmup with an analysis: every kernel loaded, the graph captured
        s.run(Wm, cluster_every=every, fpt_size=fpt_size, use_graph=use_graph)
        if args.tuning:
            fu, nc, wg = (int(v) for v in args.tuning.split(","))
            s.set_tuning(fu, nc, wg)  # recaptures the graph
            s.run(30, cluster_every=0, use_graph=use_graph)  # the steps the autotuner would take
            tuning = s.tuning()
        else:
            tuning = s.autotune(rounds=5)  # 30 more untimed steps; recaptures the graph
        log(f"[bench] {name}: tuning {tuning}")
        h0 = hamiltonian(s, params)
        nr0 = np.array([s.scalars(t).nrebuild for t in range(len(params))])
        t_all, an_ms, series, fpt = timed_run(s, K, every, use_graph, fpt_size)
        h1 = hamiltonian(s, params)
        nr1 = np.array([s.scalars(t).nrebuild for t in range(len(params))])
        sizes = np.array([p.natom for p in params])
        ev_ms = s.last_run_ms()
        tau = (K + every) * params[0].dt / W.TAU
        hdrift[name] = np.abs(h1 - h0) / np.abs(h0) / tau
        res[name] = dict(wall=t_all, ev_ms=ev_ms, an_ms=an_ms, rate=natoms * K / t_all,
                         nlist=s.list_entries(), largest=int(series.max()) if series.size else 0,
                         series=series.max(axis=1).tolist() if series.size else [],
                         nfpt=int((fpt >= 0).sum()),
                         rebuild_frac=float((sizes * (nr1 - nr0)).sum() / (natoms * (K + every))),
                         temp=temperatures(s, params), tuning=tuning)
        log(f"[bench] {name}: {K} steps + {K / every if every else 0:.2f} analyses of {an_ms:.3f} ms "
            f"-> {t_all:.4f} s (steps by events {ev_ms / 1e3:.4f} s) -> {res[name]['rate']:.4e} atom-steps/s")
        if name == "mixed":
            ferr, ferr_worst = force_error(s, params, sample)
            log(f"[bench] mixed: RMS relative force error {ferr:.3e} (worst trajectory {ferr_worst:.3e})")
            log(f"[bench] mixed: kernel timing (HIP events, {K} more steps)")
            kt = s.time_kernels(K)
            res[name]["kernels"] = kt
            res[name]["nlist_k"] = s.list_entries()
            # the drop-in hands host buffers over the boundary: one download
            # of every trajectory's x, v, a (what an output step needs)
            t0 = time.perf_counter()
            for t in range(len(params)):
                s.download(t)
            res[name]["download_s"] = time.perf_counter() - t0
        if name == "mixed_fp64_sums":
            res[name]["ferr"] = force_error(s, params, sample)
            log(f"[bench] mixed_fp64_sums: RMS relative force error {res[name]['ferr'][0]:.3e}")
        s.close()

    plain = None
    if not args.no_plain and len(params) > 1:
        # one trajectory at a time, FP64, sampled every 8th trajectory; each
        # replays its own captured 10-step graph, timed like the batch
        log("[bench] plain FP64 path, one trajectory at a time")
        kp = max(20, min(K, 100))
        tot_t = 0.0
        tot_as = 0.0
        for t in range(0, len(params), 8):
            with fcmd.System([params[t]], fcmd.FP64) as s:
                W.setup(s, [params[t]], [states[t]])
                s.run(10, cluster_every=every, use_graph=use_graph)
                dt_t, _, _, _ = timed_run(s, kp, every, use_graph, fpt_size)
                tot_t += dt_t
                tot_as += params[t].natom * kp
        plain = tot_as / tot_t
        log(f"[bench] plain FP64: {plain:.4e} atom-steps/s")

    drift_ratio = float(np.median(hdrift["mixed"]) / max(np.median(hdrift["fp64"]), 1e-300))
    t_mx, t_64 = res["mixed"]["temp"], res["fp64"]["temp"]

    # roofline: the dominant kernel of the mixed run, algorithmic bytes per
    # atom per launch (x_i, v_i, a_i write, list entries; neighbour gathers
    # are cache traffic and not counted)
    kt = res["mixed"]["kernels"]
    nl = res["mixed"]["nlist_k"] / natoms
    per_atom = {"drift": 24 + 24 + 12 + 24 + 24, "force": 24 + 24 + 12 + 4 + 4 * nl,
                "kick": 24 + 12 + 24, "rebuild": 0.0}
    dom = max(("drift", "force", "kick"), key=lambda k: kt[k])
    bytes_launch = per_atom[dom] * natoms
    achieved = bytes_launch / (kt[dom] * 1e-3)
    traffic = None
    tp = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.workload}.json")
    if os.path.exists(tp):
        traffic = json.load(open(tp)).get("hbm_bytes_per_launch")
    step_bytes = (120 + 4 * nl) * natoms  # whole-step algorithmic bytes (mixed)

    extra = None
    if not args.no_extra:
        extra = {
            "lj_nh_262k": measure("lj_nh_262k", (fcmd.FP64, fcmd.MIXED_KAHAN), steps=1000),
            "lj_nh_2m": measure("lj_nh_2m", (fcmd.FP64, fcmd.MIXED_KAHAN), steps=400, cluster_every=50),
            "mixture64": measure("mixture64", (fcmd.MIXED_KAHAN,), steps=400),
        }

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (reference Fortran on host cores)")
        try:
            cpu = cpu_baseline(params, states, args.cpu_cores or host_cores())
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] CPU baseline unavailable: {e}")
        if cpu:
            log(f"[bench] CPU baseline: {cpu['value']:.4e} atom-steps/s on {cpu['cores']} cores")

    value = res["mixed"]["rate"]
    out = {
        "metric": "atom-steps/s: mixed-prec batched vs plain FP64 on 1 GPU, <=1e-5 force err",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": res["mixed"]["wall"] / K * 1e3,
        "timing": {"method": "K steps timed by the host clock between two device synchronisations "
                             "(graph captured, kernels loaded), plus K/cluster_every analyses at "
                             "their HIP-event mean duration",
                   "analysis_ms": res["mixed"]["an_ms"], "analyses_counted": K / every if every else 0,
                   "steps_by_events_s": res["mixed"]["ev_ms"] / 1e3},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: seeded LJ vapour (uniform, 0.9 sigma minimum, Maxwell-Boltzmann)",
        "config": {"workload": args.workload, "trajectories": len(params), "atoms": natoms,
                   "cluster_every": every, "fpt_size": fpt_size, "parallelism": "1 GPU"},
        "fp64_batched": res["fp64"]["rate"],
        "fp64_plain": plain,
        "speedup_vs_fp64_plain": value / plain if plain else None,
        "speedup_vs_fp64_batched": value / res["fp64"]["rate"],
        "mixed_fp64_sums": {"value": res["mixed_fp64_sums"]["rate"],
                            "ms_per_step": res["mixed_fp64_sums"]["wall"] / K * 1e3,
                            "rms_rel_force_err": res["mixed_fp64_sums"]["ferr"][0],
                            "rms_rel_force_err_worst_trajectory": res["mixed_fp64_sums"]["ferr"][1],
                            "energy_drift_per_tau": float(np.median(hdrift["mixed_fp64_sums"])),
                            "temperature_K": float(res["mixed_fp64_sums"]["temp"].mean()),
                            "largest_cluster_series": res["mixed_fp64_sums"]["series"]},
        "rms_rel_force_err": ferr,
        "rms_rel_force_err_worst_trajectory": ferr_worst,
        "force_err_sample": f"{len(sample)} trajectories of the batch, on the mixed run's state",
        "energy_drift_ratio_mixed_over_fp64": drift_ratio,
        "energy_drift_per_tau": {"mixed": float(np.median(hdrift["mixed"])),
                                 "fp64": float(np.median(hdrift["fp64"]))},
        "temperature_K": {"mixed": float(t_mx.mean()), "fp64": float(t_64.mean()),
                          "target": float(np.mean([p.att for p in params])),
                          "max_rel_diff": float(np.max(np.abs(t_mx - t_64) / t_64))},
        "largest_cluster_series": {"mixed": res["mixed"]["series"], "fp64": res["fp64"]["series"]},
        "largest_cluster": res["mixed"]["largest"],
        "rebuilt_atom_fraction_per_step": res["mixed"]["rebuild_frac"],
        "state_download_ms": res["mixed"]["download_s"] * 1e3,
        "pcie_inclusive_output_every_100": natoms * K / (res["mixed"]["wall"] + K / 100 * res["mixed"]["download_s"]),
        "tuning": {"mixed": res["mixed"]["tuning"], "fp64": res["fp64"]["tuning"]},
        "trajectories_reaching_fpt": res["mixed"]["nfpt"],
        "kernel_ms_per_step": kt,
        "step_roofline_frac": step_bytes / (res["mixed"]["wall"] / K) / HBM_PEAK,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "bytes_per_atom": per_atom[dom], "list_entries_per_atom": nl},
        "cpu_baseline": cpu,
        "other_workloads": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
