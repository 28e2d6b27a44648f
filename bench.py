#!/usr/bin/env python
"""Throughput of the particle path on one MI355X: particle-steps/s (push + deposit).

One step = one fused LSRK4 push (5 stages, relocalization and kills after each
stage) followed by one rho + j shape-function deposition, on the cylinder
workload: 200,910 P3 tetrahedra (Np=20), 4M electrons, R=0.08, alpha=2, static
B_z=0.5 T and E_z=1e6 V/m, dt=0.2*h_min/c.  Inputs are resident in HBM before
the timed region.  Both forms run: FP64 (the drop-in) and FP32C (FP32
deposition products and interpolation, FP64 node sums); the headline value is FP32C
when its rho error against FP64 is <= 1e-6 (else FP64).  Prints one JSON line:
metric, value (particle-steps/s), ms_per_step, roofline of the dominant kernel,
the FP64 figures, the adjoint and dense-tile legs and a CPU baseline.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
METRIC = ("particle-steps/s (push+deposit), FP32-comp vs FP64 at rel ρ err≤1e-6; "
          "HBM GB/s")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_state(wl, dm, device, torch):
    from pyrticle_amd import device as dv
    x = torch.from_numpy(wl.x).to(device)
    elem = dm.find_containing_element(x)
    el_np = dv.element_numpy(elem)
    assert np.all(el_np != dv.INVALID_ELEMENT), "beam particle outside the mesh"
    # the beam is loaded element by element (stable order within an element)
    order = np.argsort(el_np, kind="stable")
    st = {
        "x": torch.from_numpy(np.ascontiguousarray(wl.x[order])).to(device),
        "p": torch.from_numpy(np.ascontiguousarray(wl.p[order])).to(device),
        "q": torch.from_numpy(np.ascontiguousarray(wl.q[order])).to(device),
        "m": torch.from_numpy(np.ascontiguousarray(wl.m[order])).to(device),
        "elem": dv.element_tensor(el_np[order], device),
        "n": wl.x.shape[0],
    }
    return st, order


def clone_state(st):
    out = {k: (v.clone() if hasattr(v, "clone") else v) for k, v in st.items()}
    return out


def run_steps(dm, wl, st, nsteps, precision, normalizer, bufs, torch):
    """nsteps of push + deposit; returns the sum of alive particles per step."""
    from pyrticle_amd import device as dv
    alive = 0
    for _ in range(nsteps):
        n = dm.lsrk4_steps(1, wl.dt, wl.c, st["n"], st["x"], st["p"], st["q"], st["m"],
                           st["elem"], precision=precision)
        st["n"] = n
        v = dv.velocities(st["p"][:n], st["m"][:n], wl.c)
        dm.deposit(wl.radius, wl.alpha, normalizer, st["x"][:n], st["q"][:n],
                   st["elem"][:n], v=v, precision=precision, rho=bufs["rho"], j=bufs["j"])
        alive += n
    return alive


def host_info():
    """Host CPU description for the CPU-baseline record."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)),
            "cpu_model": model}


def sample_state(st0, sample):
    """Every k-th particle of the element-sorted beam, as host arrays."""
    from pyrticle_amd import device as dv
    stride = max(1, st0["n"] // sample)
    sl = slice(0, stride * sample, stride)
    out = {k: np.ascontiguousarray(st0[k][sl].cpu().numpy()) for k in ("x", "p", "q", "m")}
    out["el"] = np.ascontiguousarray(dv.element_numpy(st0["elem"][sl]))
    out["stride"] = stride
    return out


def cpu_baseline(wl, smp, steps):
    """Oracle (single-threaded FP64 C restatement) on a bounded sample, on one
    pinned core: `steps` times one LSRK4 step + one rho/j deposition.  Returns
    the rate, the seconds, the reference face walk's node visits per particle
    and the final state (x, p, element ids, particle ids, rho, j) for the
    comparison with the GPU on the same sample.

    The sample starts from the same element ids as the GPU state (the
    GPU's global search is bit-exact with the oracle's O(K) scan, which would
    dominate the CPU time here)."""
    from oracle import oracle as O
    om = O.OracleMesh(wl.ma)
    x, p, q, m, el = (smp[k].copy() for k in ("x", "p", "q", "m", "el"))
    ids = np.arange(x.shape[0], dtype=np.int64)
    keep = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(keep)})       # this thread only: one core
    try:
        t0 = time.perf_counter()
        alive = 0
        visits = rho = j = None
        for _ in range(steps):
            n, _ = O.lsrk4_run(om, wl.fields, wl.c, x, p, q, m, el, wl.dt, 1, ids=ids)
            x, p, q, m, el, ids = x[:n], p[:n], q[:n], m[:n], el[:n], ids[:n]
            v = O.velocities(p, m, wl.c)
            rho, j, visits = O.deposit(om, wl.radius, wl.alpha, x, q, v, el)
            alive += n
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, keep)
    final = {"x": x, "p": p, "el": el, "ids": ids, "rho": rho, "j": j, "n": n}
    return alive / dt, dt, int(visits[0]) / max(1, n), final


def gpu_on_sample(wl, dm, smp, steps, normalizer, final, device, torch):
    """The CPU baseline's sample through the HIP path (FP64 and FP32C), the
    same steps; relative L2 errors of x, p, rho, j against the oracle's final
    state, and whether ids and element ids agree bit for bit."""
    from pyrticle_amd import device as dv

    def rel(a, b):
        nb = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))

    res = {}
    for name, prec in (("fp64", dv.FP64), ("fp32c", dv.FP32C)):
        t = {k: torch.from_numpy(smp[k].copy()).to(device) for k in ("x", "p", "q", "m")}
        el = dv.element_tensor(smp["el"], device)
        ids = torch.arange(smp["x"].shape[0], dtype=torch.int64, device=device)
        n = smp["x"].shape[0]
        rho = j = None
        for _ in range(steps):
            n = dm.lsrk4_steps(1, wl.dt, wl.c, n, t["x"], t["p"], t["q"], t["m"], el, ids=ids,
                               precision=prec)
            v = dv.velocities(t["p"][:n], t["m"][:n], wl.c)
            rho, j = dm.deposit(wl.radius, wl.alpha, normalizer, t["x"][:n], t["q"][:n], el[:n],
                                v=v, precision=prec)
        same = n == final["n"]
        r = {"alive": n, "ids_bit_exact": bool(same and np.array_equal(
                 ids[:n].cpu().numpy(), final["ids"])),
             "element_ids_bit_exact": bool(same and np.array_equal(
                 dv.element_numpy(el[:n]), final["el"])),
             "rho_rel_l2": rel(rho.cpu().numpy(), final["rho"]),
             "j_rel_l2": rel(j.cpu().numpy(), final["j"])}
        if same:
            r["x_rel_l2"] = rel(t["x"][:n].cpu().numpy(), final["x"])
            r["p_rel_l2"] = rel(t["p"][:n].cpu().numpy(), final["p"])
        res[name] = r
    return res


def square_leg(device, torch, steps=None):
    """The 2D workload (2,048 P3 triangles, 10k electrons, R=0.1, 100 steps)
    in FP64 and in its FP32-compensated form on the GPU, against the oracle's
    FP64 run on the host: 100 LSRK4 steps then one rho/j deposition.  Reports
    the GPU times, the oracle's time on one pinned core, and the relative L2
    errors of the final x, p and of rho, j against the oracle."""
    from oracle import oracle as O
    from pyrticle_amd import device as dv
    from pyrticle_amd.workloads import square_2d
    wl = square_2d()
    steps = steps or wl.nsteps
    om = O.OracleMesh(wl.ma)
    el = np.array([O.find_containing_element(om, xi) for xi in wl.x], dtype=np.uint32)
    n0 = wl.x.shape[0]
    ox, op, oq, omass, oel = wl.x.copy(), wl.p.copy(), wl.q.copy(), wl.m.copy(), el.copy()
    keep = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(keep)})
    try:
        t0 = time.perf_counter()
        n_o, _ = O.lsrk4_run(om, wl.fields, wl.c, ox, op, oq, omass, oel, wl.dt, steps)
        v_o = O.velocities(op[:n_o], omass[:n_o], wl.c)
        rho_o, j_o, _ = O.deposit(om, wl.radius, wl.alpha, ox[:n_o], oq[:n_o], v_o, oel[:n_o])
        cpu_s = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, keep)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    dm = dv.DeviceMesh(wl.ma)
    dm.set_static_fields([None if f is None else torch.from_numpy(f).to(device)
                          for f in wl.fields])
    nrm = O.shape_normalizer(wl.radius, 2, wl.alpha)
    res = {"workload": "square_2d: %d P3 triangles (Np=10), %d electrons, gamma=2, R=0.1, "
                       "alpha=2, E=(0,1e3,0) V/m, B=(0,0,0.01) T, %d LSRK4 steps then one "
                       "rho/j deposition" % (wl.ma.K, n0, steps),
           "cpu_oracle_seconds_1core": cpu_s,
           "cpu_oracle_particle_steps_per_s": n0 * steps / cpu_s}
    for name, prec in (("fp64", dv.FP64), ("fp32c", dv.FP32C)):
        best = None
        for _ in range(2):      # the first run builds the candidate tables
            t = {k: torch.from_numpy(getattr(wl, k).copy()).to(device)
                 for k in ("x", "p", "q", "m")}
            tel = dv.element_tensor(el, device)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            n = dm.lsrk4_steps(steps, wl.dt, wl.c, n0, t["x"], t["p"], t["q"], t["m"], tel,
                               precision=prec)
            v = dv.velocities(t["p"][:n], t["m"][:n], wl.c)
            rho, j = dm.deposit(wl.radius, wl.alpha, nrm, t["x"][:n], t["q"][:n], tel[:n], v=v,
                                precision=prec)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        r = {"gpu_seconds": best, "alive": n, "element_ids_bit_exact": bool(
                 n == n_o and np.array_equal(dv.element_numpy(tel[:n]), oel[:n_o])),
             "rho_rel_l2_vs_oracle": rel(rho.cpu().numpy(), rho_o),
             "j_rel_l2_vs_oracle": rel(j.cpu().numpy(), j_o)}
        if n == n_o:
            r["x_rel_l2_vs_oracle"] = rel(t["x"][:n].cpu().numpy(), ox[:n_o])
            r["p_rel_l2_vs_oracle"] = rel(t["p"][:n].cpu().numpy(), op[:n_o])
        res[name] = r
        log("square %s: %.1f ms for %d steps + deposit; x err %.2e p err %.2e rho err %.2e"
            % (name, 1e3 * best, steps, r.get("x_rel_l2_vs_oracle", -1),
               r.get("p_rel_l2_vs_oracle", -1), r["rho_rel_l2_vs_oracle"]))
    return res


def cpu_baseline_threads(wl, smp, steps, threads):
    """The same bounded sample on all host cores: particle chunks stepped and
    deposited by the oracle in parallel threads (ctypes releases the GIL),
    per-thread rho/j buffers summed afterwards.  Particles are independent in
    static fields, so the chunks need no communication."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    om = O.OracleMesh(wl.ma)
    arrays = [smp[k] for k in ("x", "p", "q", "m")]
    el = smp["el"]
    bounds = np.linspace(0, arrays[0].shape[0], threads + 1).astype(int)
    chunks = [[a[lo:hi].copy() for a in arrays] + [el[lo:hi].copy()]
              for lo, hi in zip(bounds[:-1], bounds[1:])]

    def work(ch):
        x, p, q, m, e = ch
        alive = 0
        rho = j = None
        for _ in range(steps):
            n, _ = O.lsrk4_run(om, wl.fields, wl.c, x, p, q, m, e, wl.dt, 1)
            x, p, q, m, e = x[:n], p[:n], q[:n], m[:n], e[:n]
            v = O.velocities(p, m, wl.c)
            rho, j, _ = O.deposit(om, wl.radius, wl.alpha, x, q, v, e)
            alive += n
        return alive, rho, j

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        results = list(ex.map(work, chunks))
    rho = sum(r[1] for r in results)
    dt = time.perf_counter() - t0
    del rho
    return sum(r[0] for r in results) / dt, dt


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed counter profile
    (FETCH_SIZE and WRITE_SIZE of rocprofv3 --pmc runs, each scaled by what the
    counter reads for a 512 MiB copy), or None if no profile is present."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic_cylinder3d_4M.json")
    try:
        with open(path) as f:
            return float(json.load(f)["kernels"][kernel]["hbm_bytes"])
    except (OSError, KeyError, ValueError):
        return None


# ---- rocprofv3 summaries (CPU only: reads the CSV files of profiled runs) ----
# The profiled command and the layout of its output directory:
#   <dir>/kt/run_kernel_stats.csv                 --kernel-trace --stats
#   <dir>/pmc_<C>/run_counter_collection.csv      --pmc C, one run per C below
#   <dir>/copy_<C>/run_counter_collection.csv     --pmc C over --copy-probe
PROFILE_ARGS = ("--steps 2 --warmup 1 --no-cpu-baseline --no-adjoint --no-dense --no-square "
                "--extension-reps 2 --average-steps 1")
PROFILE_COUNTERS = ("FETCH_SIZE", "WRITE_SIZE", "TCC_HIT_sum TCC_MISS_sum", "VALUBusy",
                    "MfmaUtil", "OccupancyPercent")
# traffic keys read by measured_traffic: the kernel (regular expression on the
# dispatched name; the instantiation with the most dispatches is taken)
TRAFFIC_KERNELS = {
    "k_gather_cl": r"k_gather_cl<", "k_cull_chunks": r"k_cull_chunks<",
    "k_stage32c": r"k_stage32c\b", "k_combine_items": r"k_combine_items\b",
    "k_gather_mf64": r"k_gather_mf64<", "k_stage64c": r"k_stage64c\b", "k_gather": r"k_gather<",
    "k_lsrk4_step": r"k_lsrk4_step_c<3, 3, double",
    "k_lsrk4_step_fp32c": r"k_lsrk4_step_c<3, 3, float, false", "k_stage": r"k_stage\(",
    "k_rs_scatter": r"k_rs_scatter\b", "k_shape_integral": r"k_shape_integral<",
    "k_shape_average": r"k_shape_average<", "k_reduce": r"k_reduce\b",
}


def _profile_cmd(kind, counters=None, probe=False):
    base = "cd /tmp && export TMPDIR=/tmp && rocprofv3 "
    if kind == "kt":
        return (base + "--kernel-trace --stats --output-format csv -d "
                "$GRAFT_REPO_ROOT/gpurun_out/pf/kt -o run -- python3 $GRAFT_REPO_ROOT/bench.py "
                + PROFILE_ARGS)
    tag = ("copy_" if probe else "pmc_") + counters.split()[0].replace("_sum", "")
    return (base + "--pmc %s --output-format csv -d $GRAFT_REPO_ROOT/gpurun_out/pf/%s -o run "
            "-- python3 $GRAFT_REPO_ROOT/bench.py %s"
            % (counters, tag, "--copy-probe" if probe else PROFILE_ARGS))


def _kernel_key(name):
    """Dispatched kernel name without its argument list."""
    name = name.replace("(anonymous namespace)", "{anon}")
    depth = 0
    for k, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:k]
    return name


def _counter_means(path):
    """{kernel: {"dispatches": n, counter: mean per dispatch, "grid": max grid}}."""
    import csv
    acc = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            k = _kernel_key(row["Kernel_Name"])
            d = acc.setdefault(k, {"_ids": set(), "grid": 0})
            d["_ids"].add(row["Dispatch_Id"])
            d["grid"] = max(d["grid"], int(row["Grid_Size"]))
            c = row["Counter_Name"]
            d.setdefault(c, []).append(float(row["Counter_Value"]))
    out = {}
    for k, d in acc.items():
        o = {"dispatches": len(d.pop("_ids")), "grid": d.pop("grid")}
        for c, vals in d.items():
            o[c] = sum(vals) / len(vals)
        out[k] = o
    return out


def _pick(table, pattern):
    import re
    hits = [(v["dispatches"] if isinstance(v, dict) and "dispatches" in v else v.get("calls", 0),
             k) for k, v in table.items() if re.search(pattern, k)]
    return max(hits)[1] if hits else None


def summarize_profiles(root, workload, out_json):
    """Kernel-trace table, HBM bytes per launch and counter averages from the
    CSV output of the rocprofv3 runs named by _profile_cmd; writes out_json
    (read by measured_traffic) and prints markdown tables."""
    import csv
    stats = {}
    with open(os.path.join(root, "kt", "run_kernel_stats.csv"), newline="") as f:
        for row in csv.DictReader(f):
            stats[_kernel_key(row["Name"])] = {"calls": int(row["Calls"]),
                                               "avg_us": float(row["AverageNs"]) / 1e3,
                                               "total_us": float(row["TotalDurationNs"]) / 1e3,
                                               "pct": float(row["Percentage"])}
    pmc = {}
    for c in PROFILE_COUNTERS:
        tag = c.split()[0].replace("_sum", "")
        path = os.path.join(root, "pmc_" + tag, "run_counter_collection.csv")
        if os.path.exists(path):
            for k, v in _counter_means(path).items():
                pmc.setdefault(k, {"dispatches": 0})
                pmc[k]["dispatches"] = max(pmc[k]["dispatches"], v.pop("dispatches"))
                v.pop("grid")
                pmc[k].update(v)
    cal = {"kernel": "__amd_rocclr_copyBuffer (512 MiB device-to-device copy)",
           "bytes_read": 512 << 20, "bytes_written": 512 << 20}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = _counter_means_rows(os.path.join(root, "copy_" + c, "run_counter_collection.csv"),
                                   c)
        cal[c + "_KiB"] = rows
        cal[("fetch" if c == "FETCH_SIZE" else "write") + "_factor"] = \
            (512 << 20) / (rows * 1024.0)
    kernels = {}
    for key, pat in TRAFFIC_KERNELS.items():
        name = _pick(pmc, pat)
        if name is None or "FETCH_SIZE" not in pmc[name] or "WRITE_SIZE" not in pmc[name]:
            continue
        f, w = pmc[name]["FETCH_SIZE"], pmc[name]["WRITE_SIZE"]
        kernels[key] = {"name": name, "FETCH_SIZE_KiB": f, "WRITE_SIZE_KiB": w,
                        "hbm_bytes": 1024.0 * (f * cal["fetch_factor"]
                                               + w * cal["write_factor"])}
    doc = {"workload": workload,
           "counter_unit": "KiB per dispatch (average over the run's dispatches of the kernel)",
           "calibration": cal,
           "commands": [_profile_cmd("pmc", "FETCH_SIZE"), _profile_cmd("pmc", "WRITE_SIZE"),
                        _profile_cmd("pmc", "FETCH_SIZE", probe=True),
                        _profile_cmd("pmc", "WRITE_SIZE", probe=True)],
           "kernels": kernels}
    with open(out_json, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print("## kernel trace\n\n| kernel | calls | avg µs | total µs | % |\n|---|---:|---:|---:|---:|")
    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["total_us"])[:30]:
        print("| `%s` | %d | %.1f | %.1f | %.1f |" % (k[:100], v["calls"], v["avg_us"],
                                                   v["total_us"], v["pct"]))
    print("\n## counters\n\n| kernel | avg µs (trace) | HBM GB/launch | L2 hit % | VALUBusy % "
          "| MfmaUtil % | OccupancyPercent |\n|---|---:|---:|---:|---:|---:|---:|")
    for key, pat in TRAFFIC_KERNELS.items():
        name = _pick(pmc, pat)
        if name is None:
            continue
        v = pmc[name]
        hit, miss = v.get("TCC_HIT_sum"), v.get("TCC_MISS_sum")
        l2 = 100.0 * hit / (hit + miss) if hit is not None and miss and hit + miss else None
        tr = kernels.get(key, {}).get("hbm_bytes")

        def fmt(x, spec="%.1f"):
            return spec % x if x is not None else "-"
        print("| `%s` | %s | %s | %s | %s | %s | %s |" % (
            name, fmt(stats.get(name, {}).get("avg_us")), fmt(tr / 1e9 if tr else None, "%.2f"),
            fmt(l2), fmt(v.get("VALUBusy")), fmt(v.get("MfmaUtil")),
            fmt(v.get("OccupancyPercent"))))
    print("\ncommands:")
    print(_profile_cmd("kt"))
    for c in PROFILE_COUNTERS:
        print(_profile_cmd("pmc", c))
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        print(_profile_cmd("pmc", c, probe=True))


def _counter_means_rows(path, counter):
    """Mean counter value over the copy kernel's dispatches that read at least
    half the largest value (the 512 MiB copies of --copy-probe)."""
    import csv
    vals = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter and "copyBuffer" in row["Kernel_Name"]:
                vals.append(float(row["Counter_Value"]))
    if not vals:
        raise ValueError("%s: no copy dispatch with counter %s" % (path, counter))
    top = [v for v in vals if v >= 0.5 * max(vals)]
    return sum(top) / len(top)


def adjoint_leg(wl, dm, device, torch, normalizer, n, nsteps, ckpt, fd_n):
    """Gradient of the density-matching loss w.r.t. the 6 beam parameters and
    all initial phase-space coordinates: forward = nsteps push steps + final
    deposition, backward = deposition adjoint + push adjoint with checkpoints
    every ckpt steps.  FP32C and FP64 at full size; central finite differences
    (FP64, relative step 1e-4) on a fd_n-particle beam that stays in the mesh."""
    from pyrticle_amd import adjoint as A
    from pyrticle_amd import device as dv
    from pyrticle_amd.units import SIUnitsWithNaturalConstants as U
    fields = [torch.from_numpy(f).to(device) for f in wl.fields]

    def problem(nn, z_cut):
        q = torch.full((nn,), -U.EL_CHARGE, dtype=torch.float64, device=device)
        m = torch.full((nn,), U.EL_MASS, dtype=torch.float64, device=device)
        noise = A.beam_noise(nn, 1, z_cut=z_cut, device=device)
        tnoise = A.beam_noise(nn, 2, z_cut=z_cut, device=device)

        def cfg_for(x0, prec):
            return A.PushDepositConfig(
                dmesh=dm, fields=fields, q=q, m=m,
                elem0=dm.find_containing_element(x0.detach().contiguous()), nsteps=nsteps,
                dt=wl.dt, c=wl.c, radius=wl.radius, alpha=wl.alpha, normalizer=normalizer,
                precision=prec, ckpt_every=ckpt)

        theta0 = A.default_beam_theta(device)
        theta_t = theta0.clone()
        theta_t[0] *= 1.1
        theta_t[1] *= 1.1
        with torch.no_grad():
            xt, pt = A.beam_state(theta_t, tnoise)
            rho_t, _ = A.push_and_deposit(xt.contiguous(), pt.contiguous(),
                                          cfg_for(xt, dv.FP64))
        norm_t = (rho_t * rho_t).sum()

        def loss(theta, prec, keep=None):
            x0, p0 = A.beam_state(theta, noise)
            x0 = x0.contiguous()
            p0 = p0.contiguous()
            if keep is not None:
                x0.retain_grad()
                p0.retain_grad()
                keep["x0"], keep["p0"] = x0, p0
            cfg = cfg_for(x0, prec)
            rho, _ = A.push_and_deposit(x0, p0, cfg)
            if keep is not None:
                keep["alive"] = cfg.stats["alive"]
            return ((rho - rho_t) ** 2).sum() / norm_t

        return theta0, loss

    res = {"workload": "%d particles, %d LSRK4 push steps then one rho deposition, "
                       "checkpoints every %d steps; loss ||rho_T - rho*||^2/||rho*||^2 against "
                       "a seed-2 beam with 1.1x semi-axes; 'alive' = particles still in the "
                       "mesh at the deposition" % (n, nsteps, ckpt)}
    theta0, loss = problem(n, 4.5)
    grads = {}
    for name, prec in (("fp32c", dv.FP32C), ("fp64", dv.FP64)):
        for rep in range(2):   # the first pass warms up allocations and tables
            th = theta0.clone().requires_grad_(True)
            keep = {}
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L = loss(th, prec, keep)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            L.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
        grads[name] = (th.grad.detach().clone(), keep["x0"].grad.detach().clone(),
                       keep["p0"].grad.detach().clone())
        res[name] = {"loss": float(L.detach()), "alive": keep["alive"],
                     "forward_ms": 1e3 * (t1 - t0), "backward_ms": 1e3 * (t2 - t1),
                     "grad_theta": [float(v) for v in grads[name][0].cpu()]}
        log("adjoint %s: forward %.1f ms, backward %.1f ms" % (name, 1e3 * (t1 - t0),
                                                              1e3 * (t2 - t1)))

    def rel(a, b):
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

    res["fp32c_vs_fp64_rel_l2"] = {"theta": rel(grads["fp32c"][0], grads["fp64"][0]),
                                   "x0": rel(grads["fp32c"][1], grads["fp64"][1]),
                                   "p0": rel(grads["fp32c"][2], grads["fp64"][2])}
    # finite differences on a beam subset that stays inside the mesh
    theta0, loss = problem(fd_n, 1.5)
    th = theta0.clone().requires_grad_(True)
    loss(th, dv.FP64).backward()
    g = th.grad.detach().cpu().numpy()
    fd = np.zeros(6)
    with torch.no_grad():
        for k in range(6):
            h = 1e-4 * float(theta0[k])
            tp, tm = theta0.clone(), theta0.clone()
            tp[k] += h
            tm[k] -= h
            fd[k] = (float(loss(tp, dv.FP64)) - float(loss(tm, dv.FP64))) / (2 * h)
    res["finite_differences"] = {
        "particles": fd_n, "h_relative": 1e-4, "grad_adjoint": g.tolist(), "grad_fd": fd.tolist(),
        "max_rel_err": float(np.max(np.abs(fd - g)) / np.max(np.abs(g)))}
    log("adjoint fd check: max rel err %.2e" % res["finite_differences"]["max_rel_err"])
    return res


def dense_p4_leg(device, torch, steps, warmup, n_particles):
    """Dense-tile interpolation leg: 49,896 P4 tets (Np=35), 16M electrons uniform
    in the cylinder (about 320 per element), element-sorted.  Times the fused
    LSRK4 push (5 stages, interpolation of 6 fields per stage) in FP64, in FP32C
    with the VALU contraction and in FP32C with the matrix-core contraction
    (full waves inside one element), and reports the FP32C errors against FP64."""
    from pyrticle_amd import device as dv
    from pyrticle_amd import workloads as W
    t0 = time.perf_counter()
    wl = W.cylinder_dense_p4(n_particles=n_particles)
    dm = dv.DeviceMesh(wl.ma)
    dm.set_static_fields([torch.from_numpy(f).to(device) for f in wl.fields])
    st0, _ = build_state(wl, dm, device, torch)
    occupied = int(torch.unique(st0["elem"]).numel())
    log("dense P4 setup %.1fs: K=%d Np=%d n=%d, %.0f particles per occupied element"
        % (time.perf_counter() - t0, wl.ma.K, wl.ma.Np, st0["n"], st0["n"] / occupied))
    res = {"workload": "cylinder_dense_p4: %d P4 tets (Np=35), %d electrons uniform in the "
                       "cylinder, gamma=10, Bz=0.5 T, Ez=1e6 V/m, dt=0.2*h_min/c; fused LSRK4 "
                       "push only" % (wl.ma.K, st0["n"]),
           "particles_per_occupied_element": st0["n"] / occupied, "steps": steps}
    finals = {}
    try:
        for name, prec, mf in (("fp64", dv.FP64, False), ("fp32c_valu", dv.FP32C, False),
                               ("fp32c_mfma", dv.FP32C, True)):
            dm.set_push_mfma(mf)
            if mf:   # untimed pass: share of particle-stages on the matrix cores
                dm.set_push_mfma(True, count=True)
                dm.push_mfma_waves()
                st = clone_state(st0)
                n_b = st["n"]
                st["n"] = dm.lsrk4_steps(1, wl.dt, wl.c, st["n"], st["x"], st["p"], st["q"],
                                         st["m"], st["elem"], precision=prec)
                res["mfma_share_of_particle_stages"] = dm.push_mfma_waves() / (5.0 * n_b)
                dm.set_push_mfma(True)
            st = clone_state(st0)
            for _ in range(warmup):
                st["n"] = dm.lsrk4_steps(1, wl.dt, wl.c, st["n"], st["x"], st["p"], st["q"],
                                         st["m"], st["elem"], precision=prec)
            st = clone_state(st0)
            torch.cuda.synchronize()
            dv.profile_enable(True)
            t1 = time.perf_counter()
            alive = 0
            for _ in range(steps):
                st["n"] = dm.lsrk4_steps(1, wl.dt, wl.c, st["n"], st["x"], st["p"], st["q"],
                                         st["m"], st["elem"], precision=prec)
                alive += st["n"]
            torch.cuda.synchronize()
            wall = time.perf_counter() - t1
            push_ms, push_n = dv.profile_read("push")
            dv.profile_enable(False)
            res[name] = {"particle_steps_per_s": alive / wall,
                         "ms_per_step": 1e3 * wall / steps,
                         "push_kernel_ms": push_ms / max(1, push_n)}
            finals[name] = st
            log("dense P4 %s: %.3f ms/step, push kernel %.3f ms" % (
                name, 1e3 * wall / steps, push_ms / max(1, push_n)))
    finally:
        dm.set_push_mfma(False)
    ref = finals["fp64"]
    for name in ("fp32c_valu", "fp32c_mfma"):
        st = finals[name]
        if st["n"] == ref["n"]:
            n = st["n"]
            dp_ref = ref["p"][:n] - st0["p"][:n]
            res[name]["x_rel_l2_vs_fp64"] = float(torch.linalg.norm(st["x"][:n] - ref["x"][:n])
                                                  / torch.linalg.norm(ref["x"][:n]))
            res[name]["dp_rel_l2_vs_fp64"] = float(torch.linalg.norm(
                (st["p"][:n] - st0["p"][:n]) - dp_ref) / torch.linalg.norm(dp_ref))
    res["mfma_speedup_vs_valu"] = (res["fp32c_mfma"]["particle_steps_per_s"]
                                   / res["fp32c_valu"]["particle_steps_per_s"])
    # the contraction the matrix cores take: 6 fields x 35 monomials per particle-stage
    contraction = 2 * 6 * 35 * 5 * alive / steps
    res["contraction_tflops_mfma"] = contraction / (res["fp32c_mfma"]["push_kernel_ms"] * 1e-3) / 1e12
    return res


def extensions_leg(wl, dm, device, torch, st, normalizer, reps, avg_steps, pairs_passed):
    """The components beside the headline path, on the same 4M-particle state:
    the normalized deposition (per-particle shape integrals, FP64, then the
    deposit with the charges q / I), the averaging pusher's shape-weighted force
    gather and its LSRK4 step, and the diagnostics reductions."""
    from pyrticle_amd import device as dv
    n = st["n"]
    x, q, el, p, m = (st[k][:n] for k in ("x", "q", "elem", "p", "m"))
    v = dv.velocities(p, m, wl.c)
    w = wl.discr.element_groups[0].local_discretization.integral_weights()
    Np = wl.ma.Np
    res = {}

    def timed(fn, region=None):
        fn()
        torch.cuda.synchronize()
        if region:
            dv.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        kernel = None
        if region:
            ms, cnt = dv.profile_read(region)
            dv.profile_enable(False)
            kernel = ms / max(1, cnt)
        return out, wall, kernel

    # normalized deposition against the plain one
    norm = {}
    rho_n = {}
    for name, prec in (("fp64", dv.FP64), ("fp32c", dv.FP32C)):
        _, t_plain, _ = timed(lambda: dm.deposit(wl.radius, wl.alpha, normalizer, x, q, el, v=v,
                                                 precision=prec))
        (rho, j, nz), t_norm, k_int = timed(
            lambda: dm.deposit_normalized(wl.radius, wl.alpha, normalizer, x, q, el, w, v=v,
                                          precision=prec), "integral")
        rho_n[name] = rho
        norm[name] = {"ms_per_call": 1e3 * t_norm, "plain_deposit_ms_per_call": 1e3 * t_plain,
                      "shape_integral_kernel_ms": k_int, "zero_mass_particles": nz}
    # shape integral: per passing (particle, target) pair, Np x (t: D FMAs + 1
    # sub, square, x w + add) = 8 Np flops at D = 3
    flops_int = (5.0 + wl.ma.dim) * Np * pairs_passed
    k_int = norm["fp64"]["shape_integral_kernel_ms"]
    tf_int = flops_int / (k_int * 1e-3) / 1e12 if k_int else 0.0
    norm["shape_integral_roofline"] = {
        "bound": "valu", "kernel": "k_shape_integral (FP64)", "achieved": tf_int,
        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf_int / FP64_PEAK_TFLOPS,
        "traffic": measured_traffic("k_shape_integral"), "flops_per_launch": flops_int,
        "flops_per_pair": (5 + wl.ma.dim) * Np}
    qsum = float(q.sum())
    norm["charge_rel_err_fp64"] = abs(wl.discr.integral(rho_n["fp64"].cpu().numpy())
                                      / qsum - 1)
    norm["rho_rel_l2_fp32c_vs_fp64"] = float(torch.linalg.norm(rho_n["fp32c"] - rho_n["fp64"])
                                             / torch.linalg.norm(rho_n["fp64"]))
    # the face-walk statistics pass (instrumentation of the normalized depositor)
    _, t_walk, _ = timed(lambda: dm.face_walk_stats(wl.radius, x, el))
    norm["face_walk_stats_ms_per_call"] = 1e3 * t_walk
    res["normalized_deposit"] = norm
    log("normalized deposit: fp64 %.2f ms (plain %.2f), fp32c %.2f ms (plain %.2f)"
        % (norm["fp64"]["ms_per_call"], norm["fp64"]["plain_deposit_ms_per_call"],
           norm["fp32c"]["ms_per_call"], norm["fp32c"]["plain_deposit_ms_per_call"]))

    # averaging pusher: forces, then LSRK4 steps
    fields = [torch.from_numpy(f).to(device) for f in wl.fields]
    (F, integ, nz), t_avg, k_avg = timed(
        lambda: dm.average_forces(wl.radius, wl.alpha, normalizer, fields, x, el, q, v, w),
        "average")
    # per (particle, target element) pair that passes the bound-sphere test:
    # Np nodes x (t: D FMAs + 1 sub, square, x w, + I: 10 flops) + 6 fields x
    # Np FMAs (12 flops) = 22 Np flops
    flops = 22.0 * Np * pairs_passed
    tf = flops / (k_avg * 1e-3) / 1e12 if k_avg else 0.0
    avg = {"forces_ms_per_call": 1e3 * t_avg, "gather_kernel_ms": k_avg,
           "roofline": {"bound": "valu", "kernel": "k_shape_average (FP64)", "achieved": tf,
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tf / FP64_PEAK_TFLOPS,
                        "traffic": measured_traffic("k_shape_average"),
                        "flops_per_launch": flops, "pairs_passed": pairs_passed,
                        "flops_per_pair": 22 * Np}}
    st2 = clone_state(st)
    dm.lsrk4_steps_average(1, wl.dt, wl.c, wl.radius, wl.alpha, normalizer, w, st2["n"],
                           st2["x"], st2["p"], st2["q"], st2["m"], st2["elem"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    alive = 0
    for _ in range(avg_steps):
        st2["n"] = dm.lsrk4_steps_average(1, wl.dt, wl.c, wl.radius, wl.alpha, normalizer, w,
                                          st2["n"], st2["x"], st2["p"], st2["q"], st2["m"],
                                          st2["elem"])
        alive += st2["n"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    avg["lsrk4_step"] = {"particle_steps_per_s": alive / wall, "ms_per_step": 1e3 * wall /
                         avg_steps, "steps": avg_steps}
    res["average_pusher"] = avg
    log("average pusher: forces %.2f ms (gather %.2f ms, %.2f TFLOP/s), step %.2f ms"
        % (1e3 * t_avg, k_avg or 0, tf, 1e3 * wall / avg_steps))

    # diagnostics (each call: one reduction pass over the state, host result)
    diag = {}
    for name, fn, row_bytes in (
            ("rms_energy_spread", lambda: dv.rms_energy_spread(p, m, wl.c), 32),
            ("rms_beam_emittance", lambda: dv.rms_beam_emittance(x, p, 0, 2), 48),
            ("total_charge", lambda: dv.total_charge(q), 8)):
        _, wall_d, k_d = timed(fn, "diagnostics")
        gbs = n * row_bytes / (k_d * 1e-3) / 1e9 if k_d else 0.0
        diag[name] = {"ms_per_call": 1e3 * wall_d, "kernel_ms": k_d, "bytes_per_particle":
                      row_bytes, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
    diag["traffic_k_reduce"] = measured_traffic("k_reduce")
    res["diagnostics"] = diag
    log("diagnostics: energy spread %.3f ms kernel (%.0f GB/s)"
        % (diag["rms_energy_spread"]["kernel_ms"] or 0, diag["rms_energy_spread"]["hbm_gbs"]))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--particles", type=int, default=4_000_000)
    ap.add_argument("--cpu-sample", type=int, default=200_000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32c", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run two 512 MiB device-to-device copies (the reference "
                         "kernel for scaling FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense P4 interpolation leg (16M particles, 49,896 tets)")
    ap.add_argument("--dense-particles", type=int, default=16_000_000)
    ap.add_argument("--dense-steps", type=int, default=5)
    ap.add_argument("--no-extensions", action="store_true",
                    help="skip the normalized-deposition, averaging-pusher and diagnostics "
                         "legs")
    ap.add_argument("--extension-reps", type=int, default=5)
    ap.add_argument("--average-steps", type=int, default=3)
    ap.add_argument("--no-square", action="store_true",
                    help="skip the 2D workload leg (FP64 and FP32C against the oracle)")
    ap.add_argument("--summarize-profiles", metavar="DIR",
                    help="CPU only: summarize the rocprofv3 CSV output under DIR (kt/, "
                         "pmc_<counter>/, copy_<counter>/) into profiles/"
                         "traffic_cylinder3d_4M.json and print the tables")
    args = ap.parse_args(argv)
    if args.summarize_profiles:
        summarize_profiles(args.summarize_profiles,
                           "cylinder_3d, 4M electrons, 200,910 P3 tets (FP64 and FP32C steps, "
                           "normalized deposition, averaging pusher, diagnostics)",
                           os.path.join(ROOT, "profiles", "traffic_cylinder3d_4M.json"))
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the particle path here is built and measured "
              "on a single MI355X; multi-GPU sharding is not part of this build.",
              file=sys.stderr)
        return 2

    import torch
    if args.copy_probe:
        a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8}), flush=True)
        return 0
    from pyrticle_amd import device as dv
    from pyrticle_amd import workloads as W
    from pyrticle_amd._internal import PolynomialShapeFunction

    device = torch.device("cuda:0")
    t_setup = time.perf_counter()
    log("building cylinder mesh and beam")
    wl = W.cylinder_3d(n_particles=args.particles)
    K, Np = wl.ma.K, wl.ma.Np
    dm = dv.DeviceMesh(wl.ma)
    dm.set_static_fields([torch.from_numpy(f).to(device) for f in wl.fields])
    st0, order = build_state(wl, dm, device, torch)
    normalizer = PolynomialShapeFunction(wl.radius, 3, wl.alpha).normalizer
    nn = K * Np
    bufs = {"rho": torch.empty(nn, dtype=torch.float64, device=device),
            "j": torch.empty((nn, 3), dtype=torch.float64, device=device)}
    log("setup %.1fs: K=%d Np=%d n=%d dt=%.3e" % (time.perf_counter() - t_setup, K, Np,
                                                   st0["n"], wl.dt))

    # ---- FP64 (the drop-in form) ------------------------------------------
    st = clone_state(st0)
    run_steps(dm, wl, st, args.warmup, dv.FP64, normalizer, bufs, torch)
    torch.cuda.synchronize()
    dv.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    alive = run_steps(dm, wl, st, args.steps, dv.FP64, normalizer, bufs, torch)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gather_ms, gather_n = dv.profile_read("gather")
    push_ms, push_n = dv.profile_read("push")
    sort_ms, sort_n = dv.profile_read("sort")
    dv.profile_enable(False)
    wall = t1 - t0
    value = alive / wall
    log("fp64: %.3f ms/step, %.3e particle-steps/s; gather %.3f ms, push %.3f ms, sort %.3f ms"
        % (1e3 * wall / args.steps, value, gather_ms / max(1, gather_n),
           push_ms / max(1, push_n), sort_ms / max(1, sort_n)))

    # work counts of the last state (counting pass, outside the timed region)
    n_last = st["n"]
    dm.set_deposit_counting(True)
    v = dv.velocities(st["p"][:n_last], st["m"][:n_last], wl.c)
    rho64, j64 = dm.deposit(wl.radius, wl.alpha, normalizer, st["x"][:n_last],
                            st["q"][:n_last], st["elem"][:n_last], v=v)
    n_eval, n_in, n_cand, n_pass = dm.deposit_stats()
    dm.set_deposit_counting(False)
    occupied = int(torch.unique(st["elem"][:n_last]).numel())

    # ---- FP32C ------------------------------------------------------------
    fp32c = None
    if not args.no_fp32c:
        rho32, j32 = dm.deposit(wl.radius, wl.alpha, normalizer, st["x"][:n_last],
                                st["q"][:n_last], st["elem"][:n_last], v=v,
                                precision=dv.FP32C)
        list_entries, list_groups, list_instr, list_items = dm.deposit_list_stats()
        rho_err = float(torch.linalg.norm(rho32 - rho64) / torch.linalg.norm(rho64))
        j_err = float(torch.linalg.norm(j32 - j64) / torch.linalg.norm(j64))
        st32 = clone_state(st0)
        run_steps(dm, wl, st32, args.warmup, dv.FP32C, normalizer, bufs, torch)
        torch.cuda.synchronize()
        dv.profile_enable(True)
        t0 = time.perf_counter()
        alive32 = run_steps(dm, wl, st32, args.steps, dv.FP32C, normalizer, bufs, torch)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g32_ms, g32_n = dv.profile_read("gather")
        p32_ms, p32_n = dv.profile_read("push")
        dv.profile_enable(False)
        n32 = min(st32["n"], st["n"])
        x_err = float(torch.linalg.norm(st32["x"][:n32] - st["x"][:n32])
                      / torch.linalg.norm(st["x"][:n32])) if st32["n"] == st["n"] else None
        p_err = float(torch.linalg.norm(st32["p"][:n32] - st["p"][:n32])
                      / torch.linalg.norm(st["p"][:n32])) if st32["n"] == st["n"] else None
        fp32c = {"value": alive32 / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / args.steps,
                 "speedup_vs_fp64": (alive32 / (t1 - t0)) / value,
                 "rho_rel_l2_vs_fp64": rho_err, "j_rel_l2_vs_fp64": j_err,
                 "x_rel_l2_vs_fp64": x_err, "p_rel_l2_vs_fp64": p_err,
                 "gather_ms": g32_ms / max(1, g32_n), "push_ms": p32_ms / max(1, p32_n),
                 "lists": {"chunk_target_entries": list_entries,
                           "chunk_group_pairs": list_groups,
                           "mfma_4x4x1_instructions": list_instr, "work_items": list_items}}
        log("fp32c: %.3f ms/step (x%.2f), rho err %.2e" % (fp32c["ms_per_step"],
                                                           fp32c["speedup_vs_fp64"], rho_err))

    # ---- normalized deposition, averaging pusher, diagnostics -------------
    extensions = None
    if not args.no_extensions:
        extensions = extensions_leg(wl, dm, device, torch, st, normalizer, args.extension_reps,
                                    args.average_steps, n_pass)

    # ---- adjoint (gradients of a density-matching loss) -------------------
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(wl, dm, device, torch, normalizer, args.particles,
                              args.adjoint_steps, 20, 10_000)
        dm.set_static_fields([torch.from_numpy(f).to(device) for f in wl.fields])

    # ---- dense P4 tiles: matrix-core interpolation ------------------------
    dense = None
    if not args.no_dense:
        dense = dense_p4_leg(device, torch, args.dense_steps, 1, args.dense_particles)

    # ---- CPU baseline, and the GPU on the same sample ----------------------
    cpu = None
    if not args.no_cpu_baseline:
        smp = sample_state(st0, args.cpu_sample)
        log("cpu baseline: oracle on %d particles x %d steps" % (args.cpu_sample,
                                                                 args.cpu_steps))
        rate, secs, sample_visits, final = cpu_baseline(wl, smp, args.cpu_steps)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": "%d particles (every %dth of the element-sorted beam) x %d steps "
                         "(LSRK4 push + rho/j deposition) on the full %d-element mesh, %.1f s "
                         "on one pinned core" % (args.cpu_sample, smp["stride"], args.cpu_steps,
                                                 K, secs),
               "host": host_info(),
               "node_visits_per_particle_ref_walk": sample_visits}
        threads = len(os.sched_getaffinity(0))
        rate_mt, secs_mt = cpu_baseline_threads(wl, smp, args.cpu_steps, threads)
        cpu["all_cores"] = {"value": rate_mt, "cores": threads, "seconds": secs_mt,
                            "note": "the same sample in %d particle chunks, one oracle "
                                    "thread each, per-thread rho/j buffers" % threads}
        log("cpu baseline: %.3e particle-steps/s on 1 core, %.3e on %d threads"
            % (rate, rate_mt, threads))
        cpu["gpu_vs_oracle_on_sample"] = gpu_on_sample(wl, dm, smp, args.cpu_steps, normalizer,
                                                       final, device, torch)
        log("gpu vs oracle on the sample: %s" % json.dumps(cpu["gpu_vs_oracle_on_sample"]))

    square = None
    if not args.no_square:
        square = square_leg(device, torch)

    # ---- rooflines ----------------------------------------------------------
    # Algorithmic work of the deposition (the reference's): 9 flops per node
    # visit of the face walk (distance test) plus 12 per node inside R (shape
    # value, charge scaling, rho and three j accumulations).  The walk's node
    # visits are counted for every particle of the final state by the GPU's
    # face-walk kernel (elements visited x Np); pairs inside R by the FP64
    # gather's counting pass.
    n_el_walk, *_ = dm.face_walk_stats(wl.radius, st["x"][:n_last], st["elem"][:n_last])
    walk_visits = int(n_el_walk.to(torch.int64).sum()) * Np
    flops = 9.0 * walk_visits + 12.0 * n_in
    traffic_note = ("HBM bytes per launch replayed from profiles/traffic_cylinder3d_4M.json "
                    "(rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this benchmark, each scaled by "
                    "the counter's reading for a 512 MiB copy); not measured in this run")
    gather_avg_s = gather_ms / max(1, gather_n) / 1e3
    achieved_tflops = flops / gather_avg_s / 1e12 if gather_avg_s > 0 else 0.0
    work = {"flops_per_launch": flops, "model": "9 flops per node visit of the reference face "
            "walk + 12 per node inside R", "walk_node_visits": walk_visits,
            "walk_node_visits_per_particle": walk_visits / max(1, n_last),
            "pairs_inside_R": n_in, "particles": n_last}
    roof64 = {"bound": "mfma", "kernel": "k_gather_mf64 (rho+j deposition over culled chunk "
              "lists, FP64 matrix cores)",
              "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
              "frac": achieved_tflops / FP64_PEAK_TFLOPS,
              "traffic": measured_traffic("k_gather_mf64"),
              "traffic_source": traffic_note, "avg_launch_ms": gather_avg_s * 1e3,
              "work": dict(work, pairs_evaluated=n_eval, candidate_particle_visits=n_cand,
                           prefilter_passed=n_pass)}
    push_avg_s = push_ms / max(1, push_n) / 1e3
    field_bytes = occupied * (6 * Np * 8 + 12 * 8 + 4 * (3 + 1) * 8 + 4 * 4)
    push_bytes = n_last * (68 + 52) + field_bytes
    push_gbs = push_bytes / push_avg_s / 1e9 if push_avg_s > 0 else 0.0
    roof_push = {"bound": "hbm",
                 "kernel": "k_lsrk4_step_c (fused 5-stage push on per-wave LDS element slots, FP64)",
                 "achieved": push_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": push_gbs / HBM_PEAK_GBS, "traffic": measured_traffic("k_lsrk4_step"),
                 "traffic_source": traffic_note,
                 "avg_launch_ms": push_avg_s * 1e3, "bytes_per_launch": push_bytes,
                 "bytes_model": "68 B read + 52 B written per particle, plus the 6 x Np FP64 "
                                "field coefficients, inverse map and faces of each occupied "
                                "element once", "occupied_elements": occupied}

    fp64 = {"value": value, "ms_per_step": 1e3 * wall / args.steps, "dtype": "f64",
            "gather_ms": gather_avg_s * 1e3, "push_ms": push_avg_s * 1e3,
            "sort_ms": sort_ms / max(1, sort_n), "roofline": roof64}
    headline32 = fp32c is not None and fp32c["rho_rel_l2_vs_fp64"] <= 1e-6
    if headline32:
        # FP32C gather: the same algorithmic work against the FP32 peak (the
        # matrix cores and the VALU share it: 157.3 TFLOP/s).  The matrix-core
        # flops the list gather issues (512 per 4x4x1 instruction, counted
        # exactly from the last deposition's lists) are reported beside it.
        # The region timed as "gather" is the list gather k_gather_cl alone;
        # the fixed-order combination of its partials (k_combine_items) is
        # timed with the sort and the other set-up work.
        g32_s = fp32c["gather_ms"] / 1e3
        tflops = flops / g32_s / 1e12 if g32_s > 0 else 0.0
        issued = 512.0 * fp32c["lists"]["mfma_4x4x1_instructions"]
        roofline = {"bound": "mfma", "kernel": "k_gather_cl (rho+j deposition over culled "
                    "chunk lists, FP32 matrix cores, FP64 node sums)", "achieved": tflops,
                    "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": tflops / FP32_MFMA_PEAK_TFLOPS,
                    "traffic": measured_traffic("k_gather_cl"), "traffic_source": traffic_note,
                    "avg_launch_ms": fp32c["gather_ms"],
                    "work": dict(work, issued_mfma_flops_per_launch=issued,
                                 issued_mfma_tflops=issued / g32_s / 1e12 if g32_s > 0 else 0.0,
                                 lists=fp32c["lists"])}
    else:
        roofline = roof64

    # whole-step roofline: the step's algorithmic bytes (212 B per
    # particle-step: state read 68 + written 52, nodal fields 48, mesh 12,
    # rho/j 32) against HBM and its algorithmic flops (the deposition work
    # model above plus 1.7 kflop per particle-step for the five push stages)
    # against the vector peak of the form's precision; frac = the larger
    # bound time over the measured step time
    def step_roofline(ms_step, peak_tflops):
        n_p = n_last
        t_hbm = 212.0 * n_p / (HBM_PEAK_GBS * 1e9)
        step_flops = flops + 1700.0 * n_p
        t_flop = step_flops / (peak_tflops * 1e12)
        bound = max(t_hbm, t_flop)
        return {"bytes_per_particle_step": 212, "flops_per_step": step_flops,
                "hbm_bound_ms": 1e3 * t_hbm, "flop_bound_ms": 1e3 * t_flop,
                "peak_tflops": peak_tflops, "measured_ms": ms_step,
                "frac": 1e3 * bound / ms_step if ms_step > 0 else 0.0}

    fp64["step_roofline"] = step_roofline(fp64["ms_per_step"], FP64_PEAK_TFLOPS)
    if fp32c is not None:
        fp32c["step_roofline"] = step_roofline(fp32c["ms_per_step"], FP32_MFMA_PEAK_TFLOPS)

    out = {
        "metric": METRIC,
        "value": fp32c["value"] if headline32 else value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp32c["ms_per_step"] if headline32 else 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32c" if headline32 else "f64",
        "data": "synthetic (seeded beam and mesh; static fields)",
        "config": {"workload": "cylinder_3d: r=1, z in [0,4], %d P3 tets (Np=20), %d "
                               "electrons (KV-like ellipse 0.3x0.2, sigma_z=0.4, gamma=10, "
                               "0.1%% spread), R=0.08, alpha=2, Bz=0.5 T, Ez=1e6 V/m, "
                               "dt=0.2*h_min/c, one step = LSRK4 push + rho/j deposition"
                               % (K, st0["n"]),
                   "elements": K, "particles": st0["n"],
                   "precision": ("fp32c: FP32 interpolation, shape evaluation and deposition "
                                 "products, FP32 partial sums folded into FP64 node sums "
                                 "every 64 chunks; FP64 positions, momenta, face tests and "
                                 "charge normalization" if headline32 else "fp64")},
        "roofline": roofline,
        "roofline_push": roof_push,
        "step_roofline": fp32c["step_roofline"] if headline32 else fp64["step_roofline"],
        "fp64": fp64,
        "fp32c": fp32c,
        "square_2d": square,
        "adjoint": adjoint,
        "dense_p4": dense,
        "extensions": extensions,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
