"""Benchmark: interface cell-updates/s of 3-D isotropic grain growth on one MI355X.

Workload: 256^3 periodic grid, Voronoi labels from 20,000 sites (seed 42),
IsotropicGrowthThermo(M0 = 1, sigma0 = 1), D = 0, geometry sigma = 1.5 cells,
dt at the stable limit.  A step is one GPKGrid::step of the whole grid.

The headline `value` is the FP64 GPU path: sum over the K timed steps of the
band size |A_s| divided by the wall time of those steps, with the state
already resident in HBM.  The FP32 path is timed the same way and reported
beside it, with the curvature error against FP64 on identical labels and the
label disagreement after the run.  `full_run` times the whole configured
run (--full-steps, 200 by default) per precision from the Voronoi state,
step by step, with the band size range and, at its end, the FP32 label
disagreement and curvature error.  The CPU baseline times the reference step
(oracle/_ref, built from the reference sources) on this host's cores for a
bounded sample of the same workload: one untimed warm-up step, then
--cpu-steps timed steps (3 by default), recording the CPU model and the
OpenMP thread count.  `roofline` is the geometry kernel's executed FP64 flops
per launch over its HIP-event launch time against the FP64 vector peak (its
dense-formulation rate under `effective`), with the HBM bytes per launch
from the committed rocprofv3 counter summary.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]

`--workload batch` runs the lamellar sweep: 512 2-D instances, side
~ U_int[64, 384], stripe width t ~ U_int[6, 20], period round(r t) with
r ~ U[2, 12], one-cell seeded edge perturbations (numpy default_rng(0)),
AnisoOstwaldThermo(M0 1, sigma0 1, bxx 1, byy 0.5, bzz 1, c_sat 0.1,
Omega 1e8) at c = c_sat (pure capillary flow), D = 0, sigma 1.5; 300 steps
in FP32 and FP64 through the batched persistent kernel after the autotuner
has picked its launch shape.  value = instance-steps/s.

`--workload batch_grain` runs 512 2-D grain-growth instances (side
~ U_int[64, 256], numpy default_rng(7); one periodic Voronoi grain per 64
cells from the library's host generator), isotropic growth with the
grain-network topology test, sigma 1.5, 200 steps in FP32 and FP64 through
the same batched kernel.  value = instance-steps/s.

`--workload wide_sigma` runs 128x256x256 grain growth at sigma 10.67 cells
(240 grains), the per-label separable geometry path.

`--workload solidification` instead runs the solute-coupled case: 128^3,
SolidificationThermo(M0 5e-8, sigma0 0.2, eps4 0.02, m_liq -500, k 0.14,
T_melt 933, L 1e8) with partitioned transport (solute_rejection), D = 1e-9,
dx = 1 um, sigma 1.5, a solid sphere r = 6 at the centre (c = c0 k inside),
c = c0 + U(-1e-3, 1e-3) outside (numpy default_rng(7)), undercooling
0.5 |m| c0 (1 - k) / k; 500 steps per precision, the FP64 solute sum taken
after every step and its relative drift reported.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# CPU-baseline thread placement (one thread per core, kept close), set before
# any OpenMP runtime is loaded; the thread count is the environment's
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "strict32": 157.3}


def run_gpu(prec, size, grains, sigma, warmup, steps):
    import gpk_amd
    from gpk_amd import _abi

    model = _abi.isotropic(1.0, 1.0)
    G = gpk_amd.GPKGrid(size, size, size, 1.0, model, 0.0, sigma, precision=prec)
    t0 = time.time()
    G.voronoi_init(grains, 42)
    G.sync()
    t_init = time.time() - t0
    for _ in range(warmup):
        G.step()
    G.sync()
    G.set_profiling(True)
    t0 = time.perf_counter()
    st = G.step_many(steps)
    G.sync()
    wall = time.perf_counter() - t0
    geo_ms, launches, flops = G.kernel_timing(0)
    G.set_profiling(False)
    return G, {
        "precision": prec,
        "steps": steps,
        "warmup": warmup,
        "band_cells_total": int(st.band_cells),
        "band_fraction": float(st.band_cells) / (steps * size ** 3),
        "seconds": wall,
        "ms_per_step": 1e3 * wall / steps,
        "cell_updates_per_s": st.band_cells / wall,
        "accepted": int(st.accepted),
        "deferred": int(st.deferred),
        "exchanges": int(st.exchanges),
        "guard_band_decisions": int(st.guarded),
        "geometry_ms_avg": geo_ms,
        "geometry_launches": launches,
        "geometry_flops_per_launch": flops,
        "init_seconds": t_init,
    }


def kappa_error_on(phi, size, sigma):
    """FP32 vs FP64 curvature on identical labels: max and RMS of
    |k32 - k64| / max(|k64|, 1e-3 / (sigma dx)) over the band."""
    import gpk_amd
    from gpk_amd import _abi

    model = _abi.isotropic(1.0, 1.0)
    out = {}
    G64 = gpk_amd.GPKGrid(size, size, size, 1.0, model, 0.0, sigma, precision="fp64")
    G64.set("phi", phi)
    G64.step(0.0)
    k64 = G64.get("kappa").ravel()
    band = G64.get("mask").ravel().astype(bool)
    G64.close()
    G32 = gpk_amd.GPKGrid(size, size, size, 1.0, model, 0.0, sigma, precision="fp32")
    G32.set("phi", phi)
    G32.step(0.0)
    k32 = G32.get("kappa").ravel().astype(np.float64)
    G32.close()
    floor = 1e-3 / sigma
    rel = np.abs(k32[band] - k64[band]) / np.maximum(np.abs(k64[band]), floor)
    out = {"kappa_rel_err_max": float(rel.max()), "kappa_rel_err_rms": float(np.sqrt(np.mean(rel ** 2))),
           "kappa_floor": floor}
    return out


def run_full(prec, size, grains, sigma, steps):
    """The whole configured run (steps from the Voronoi state), one step per
    call so that every step's band size is recorded."""
    import gpk_amd
    from gpk_amd import _abi

    G = gpk_amd.GPKGrid(size, size, size, 1.0, _abi.isotropic(1.0, 1.0), 0.0, sigma, precision=prec)
    G.voronoi_init(grains, 42)
    G.sync()
    bands = []
    t0 = time.perf_counter()
    for _ in range(steps):
        G.step()
        bands.append(int(G.last_stats.band_cells))
    G.sync()
    wall = time.perf_counter() - t0
    phi = G.get("phi")
    G.close()
    b = np.array(bands, np.float64)
    return phi, {
        "precision": prec, "steps": steps, "seconds": wall, "ms_per_step": 1e3 * wall / steps,
        "cell_updates_per_s": float(b.sum()) / wall,
        "band_cells_min": int(b.min()), "band_cells_max": int(b.max()),
        "band_fraction_first": float(b[0]) / size ** 3, "band_fraction_last": float(b[-1]) / size ** 3,
    }


def band_size(phi):
    """|A|: cells with a periodic face neighbour of another label (src/gpk_geometry.cpp:42-82)."""
    m = np.zeros(phi.shape, bool)
    for ax in range(phi.ndim):
        for sh in (1, -1):
            m |= phi != np.roll(phi, sh, axis=ax)
    return int(np.count_nonzero(m))


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = None
    try:
        threads = int(C.CDLL("libgomp.so.1").omp_get_max_threads())
    except OSError:
        pass
    return model, threads


def cpu_baseline(phi, g, size, sigma, cpu_steps):
    """Reference CPU step (compiled reference sources) on this host: one
    untimed warm-up step (first touch of the reference's work arrays), then
    cpu_steps timed steps, each bracketed by a steady clock."""
    from gpk_amd import _abi
    from oracle import oracle as O

    model = _abi.isotropic(1.0, 1.0)
    kind = "reference" if O.ref_available() else "port"
    if kind == "reference":
        R = O.RefGrid(size, size, size, 1.0, model, 0.0, sigma)
    else:
        R = O.OracleGrid(size, size, size, 1.0, model, 0.0, sigma, "strict32")
    R.set("phi", phi)
    R.set("g", g)
    R.step()
    band = 0
    wall = 0.0
    for _ in range(cpu_steps):
        band += band_size(R.get("phi"))
        t0 = time.perf_counter()
        R.step()
        wall += time.perf_counter() - t0
    cpu_model, threads = host_cpu()
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": band / wall, "unit": "interface cell-updates/s", "cores": threads,
            "kind": kind, "seconds": wall, "cpu_model": cpu_model, "nproc": os.cpu_count(),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
            "sample": "reference steps 2-%d of the same %d^3 Voronoi state (step 1 untimed warm-up)"
                      % (cpu_steps + 1, size)}


def solidification_setup(N, rng_seed=7, c0=0.02, k=0.14, undercooling=0.5, r=6):
    m_liq, T_melt = -500.0, 933.0
    from gpk_amd import _abi
    model = _abi.solidification(M0=5e-8, sigma0=0.2, eps4=0.02, m_liq=m_liq, k_part=k,
                                T_melt=T_melt, L=1e8, solute_rejection=True)
    zz, yy, xx = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    ctr = N // 2
    inside = (zz - ctr) ** 2 + (yy - ctr) ** 2 + (xx - ctr) ** 2 <= r * r
    phi = inside.astype(np.int32)
    conc = (c0 + np.random.default_rng(rng_seed).uniform(-1e-3, 1e-3, size=phi.shape)).astype(np.float32)
    conc[inside] = np.float32(c0 * k)
    dT = undercooling * abs(m_liq) * c0 * (1 - k) / k
    T = np.full(phi.shape, T_melt + m_liq * c0 - dT, np.float32)
    return model, phi, conc, T


def run_solidification(prec, N, steps, sigma):
    """Partitioned-transport solidification; per-step FP64 mass and timing."""
    import gpk_amd
    model, phi, conc, T = solidification_setup(N)
    G = gpk_amd.GPKGrid(N, N, N, 1e-6, model, 1e-9, sigma, precision=prec)
    G.set("phi", phi)
    G.set("c", conc)
    G.set("T", T)
    M0 = G.solute_sum()
    drift = []
    band = 0
    step_s = 0.0
    for _ in range(steps):
        G.sync()
        t0 = time.perf_counter()
        G.step()
        G.sync()
        step_s += time.perf_counter() - t0
        band += int(G.last_stats.band_cells)
        drift.append((G.solute_sum() - M0) / M0)
    d = np.abs(np.array(drift))
    sf = G.solid_fraction()
    G.close()
    return {"precision": prec, "steps": steps, "ms_per_step": 1e3 * step_s / steps,
            "cell_updates_per_s": band / step_s, "band_cells_total": band,
            "mass_rel_drift_max": float(d.max()), "mass_rel_drift_final": float(drift[-1]),
            "solid_fraction_final": float(sf), "M0": M0}


def main_solidification(args):
    N = args.size if args.size != 256 else 128
    steps = args.steps if args.steps != 20 else 500
    res = {p: run_solidification(p, N, steps, args.sigma) for p in ("fp32", "fp64")}
    r64 = res["fp64"]
    out = {
        "metric": "interface cell-updates/s, 128^3 solute-coupled solidification",
        "value": r64["cell_updates_per_s"], "unit": "interface cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": 0, "ms_per_step": r64["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded sphere)",
        "config": {"workload": "3-D partitioned solidification, %d^3, r = 6 seed, %d steps" % (N, steps),
                   "grid": [N] * 3, "precision": "fp64", "parallelism": "single GPU"},
        "fp64": r64, "fp32": res["fp32"],
    }
    print(json.dumps(out))


def batch_model():
    from gpk_amd import _abi
    return _abi.aniso_ostwald(1.0, 1.0, 1.0, 0.5, 1.0, float(np.float32(0.1)), 1e8)


def lamellar_fields(shapes):
    return {"c": [np.full(s, np.float32(0.1), np.float32) for s in shapes]}


def run_batch(prec, shapes, phis, steps, sigma, sample_steps=20, model=None, fields=None):
    from gpk_amd.batch import GPKBatch
    B = GPKBatch(shapes, 1.0, model or batch_model(), 0.0, sigma, precision=prec)
    B.set("phi", phis)
    for name, arrays in (fields if fields is not None else lamellar_fields(shapes)).items():
        B.set(name, arrays)
    t0 = time.perf_counter()
    wg, nwg, tune_ms = B.autotune()
    t_tune = time.perf_counter() - t0
    B.step(1)  # graph capture + warm-up step
    t0 = time.perf_counter()
    B.step(steps)
    wall = time.perf_counter() - t0
    st = B.stats()
    phi = B.get("phi")
    per = []
    for _ in range(sample_steps):
        B.step(1)
        per.append(B.instances()[1])
    per = np.concatenate(per)
    B.close()
    n = len(shapes)
    return phi, {
        "precision": prec, "instances": n, "steps": steps, "seconds": wall,
        "ms_per_step": 1e3 * wall / steps, "instance_steps_per_s": n * steps / wall,
        "cell_updates_per_s": st["band_cells"] * steps / (steps + 1) / wall,
        "accepted": int(st["accepted"]), "exchanges": int(st["exchanges"]),
        "deferred": int(st["deferred"]), "eroded": int(st["eroded"]),
        "guard_band_decisions": int(st["guarded"]),
        "instance_step_ms_p50": float(np.percentile(per, 50)),
        "instance_step_ms_p99": float(np.percentile(per, 99)),
        "autotune": {"workgroup_size": wg, "workgroups": nwg, "step_ms": tune_ms, "seconds": t_tune},
    }


def cpu_batch_baseline(shapes, phis, sigma, n_inst=32, steps=30, model=None, fields=None):
    """Reference GPKGrid steps (compiled reference sources) on a sample of instances."""
    from oracle import oracle as O
    kind = "reference" if O.ref_available() else "port"
    model = model or batch_model()
    fields = fields if fields is not None else lamellar_fields(shapes)
    wall = 0.0
    for b, (s, p) in enumerate(list(zip(shapes, phis))[:n_inst]):
        if kind == "reference":
            R = O.RefGrid(1, s[0], s[1], 1.0, model, 0.0, sigma)
        else:
            R = O.OracleGrid(1, s[0], s[1], 1.0, model, 0.0, sigma, "strict32")
        R.set("phi", p)
        for name, arrays in fields.items():
            R.set(name, arrays[b])
        t0 = time.perf_counter()
        for _ in range(steps):
            R.step()
        wall += time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n_inst * steps / wall, "unit": "instance-steps/s", "cores": threads, "kind": kind,
            "seconds": wall,
            "sample": "%d reference steps of each of the first %d instances, run one after another" %
                      (steps, n_inst)}


def main_batch(args):
    from gpk_amd.batch import lamellar_sweep
    n = 512
    steps = args.steps if args.steps != 20 else 300
    shapes, phis, _ = lamellar_sweep(n, seed=0)
    p32, r32 = run_batch("fp32", shapes, phis, steps, args.sigma)
    p64, r64 = run_batch("fp64", shapes, phis, steps, args.sigma)
    diff = sum(int(np.count_nonzero(a != b)) for a, b in zip(p32, p64))
    r32["label_disagreement_vs_fp64"] = diff / float(sum(a.size for a in p32))
    cpu = None if args.no_cpu else cpu_batch_baseline(shapes, phis, args.sigma)
    out = {
        "metric": "instance-steps/s, 512 batched 2-D lamellar instances",
        "value": r32["instance_steps_per_s"], "unit": "instance-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": r32["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded lamellar sweep)",
        "config": {"workload": "512 2-D instances, side U[64,384], lamellae t U[6,20], period r t, "
                               "r U[2,12], AnisoOstwald c = c_sat, %d steps" % steps,
                   "instances": n, "cells": int(sum(a.size for a in p32)), "precision": "fp32",
                   "parallelism": "single GPU"},
        "cpu_baseline": cpu, "fp32": r32, "fp64": r64,
    }
    if cpu:
        out["speedup_vs_cpu"] = r32["instance_steps_per_s"] / cpu["value"]
    print(json.dumps(out))


def main_batch_grain(args):
    """Batched 2-D isotropic grain growth with the grain-network topology
    test: 512 instances of side U[64, 256] (seeded), each a periodic Voronoi
    map of one grain per 64 cells on average (host Voronoi generator of the
    library), orientations from the generator; flips decided one after
    another per instance, unresolved grains eroded every 8th step."""
    import gpk_amd
    from gpk_amd import _abi
    n = 512
    steps = args.steps if args.steps != 20 else 200
    rng = np.random.default_rng(7)
    shapes, phis, gs = [], [], []
    for b in range(n):
        ny, nx = (int(v) for v in rng.integers(64, 257, 2))
        ng = max(4, min(1000, ny * nx // 64))
        p, g = gpk_amd.voronoi_host(1, ny, nx, ng, 1000 + b)
        shapes.append((ny, nx))
        phis.append(p.reshape(ny, nx))
        gs.append(g.reshape(ny, nx))
    model = _abi.isotropic(1.0, 1.0)
    fields = {"g": gs}
    p32, r32 = run_batch("fp32", shapes, phis, steps, args.sigma, model=model, fields=fields)
    p64, r64 = run_batch("fp64", shapes, phis, steps, args.sigma, model=model, fields=fields)
    diff = sum(int(np.count_nonzero(a != b)) for a, b in zip(p32, p64))
    r32["label_disagreement_vs_fp64"] = diff / float(sum(a.size for a in p32))
    cpu = None if args.no_cpu else cpu_batch_baseline(shapes, phis, args.sigma, model=model, fields=fields)
    out = {
        "metric": "instance-steps/s, 512 batched 2-D grain-growth instances (grain-network topology test)",
        "value": r32["instance_steps_per_s"], "unit": "instance-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": r32["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded periodic Voronoi maps)",
        "config": {"workload": "512 2-D instances, side U[64,256], one Voronoi grain per 64 cells, "
                               "isotropic growth, sigma %.2f, %d steps" % (args.sigma, steps),
                   "instances": n, "cells": int(sum(a.size for a in p32)), "precision": "fp32",
                   "parallelism": "single GPU"},
        "cpu_baseline": cpu, "fp32": r32, "fp64": r64,
    }
    if cpu:
        out["speedup_vs_cpu"] = r32["instance_steps_per_s"] / cpu["value"]
    print(json.dumps(out))


def main_wide_sigma(args):
    """Grain growth at a wide Gaussian: 128 x 256 x 256, sigma 10.67 cells,
    240 Voronoi grains (the shape of the reference's fine small-grain 3-D
    decks).  The geometry radius is 40 > 18, so every step takes the
    per-label separable path; the projection window has radius 9."""
    import gpk_amd
    from gpk_amd import _abi
    sigma = 10.666666666667
    steps = args.steps if args.steps != 20 else 5
    res = {}
    for prec in ("fp64", "fp32"):
        G = gpk_amd.GPKGrid(128, 256, 256, 1.0, _abi.isotropic(1.0, 1.0), 0.0, sigma, precision=prec)
        G.voronoi_init(240, 42)
        G.step()
        G.sync()
        t0 = time.perf_counter()
        st = G.step_many(steps)
        G.sync()
        wall = time.perf_counter() - t0
        res[prec] = {"precision": prec, "steps": steps, "ms_per_step": 1e3 * wall / steps,
                     "cell_updates_per_s": st.band_cells / wall, "band_cells_total": int(st.band_cells),
                     "accepted": int(st.accepted), "exchanges": int(st.exchanges)}
        G.close()
    r64 = res["fp64"]
    out = {
        "metric": "interface cell-updates/s, 128x256x256 grain growth at sigma 10.67 (per-label separable geometry)",
        "value": r64["cell_updates_per_s"], "unit": "interface cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": 1, "ms_per_step": r64["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (periodic Voronoi, 240 sites, seed 42)",
        "config": {"workload": "3-D isotropic grain growth, 128x256x256, 240 grains, sigma 10.67 cells",
                   "grid": [128, 256, 256], "grains": 240, "sigma_cells": sigma, "precision": "fp64",
                   "parallelism": "single GPU"},
        "fp64": r64, "fp32": res["fp32"],
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["grain_growth", "solidification", "batch", "batch_grain",
                                           "wide_sigma"],
                    default="grain_growth")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--grains", type=int, default=20000)
    ap.add_argument("--sigma", type=float, default=1.5)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--full-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the per-step kinetics path is built and measured "
              "on a single MI355X (multi-GPU sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "solidification":
        main_solidification(args)
        return
    if args.workload == "batch":
        main_batch(args)
        return
    if args.workload == "batch_grain":
        main_batch_grain(args)
        return
    if args.workload == "wide_sigma":
        main_wide_sigma(args)
        return

    G64, r64 = run_gpu("fp64", args.size, args.grains, args.sigma, args.warmup, args.steps)
    res32 = None
    if not args.no_fp32:
        G32, res32 = run_gpu("fp32", args.size, args.grains, args.sigma, args.warmup, args.steps)
        p64 = G64.get("phi")
        p32 = G32.get("phi")
        res32["label_disagreement"] = float(np.count_nonzero(p64 != p32)) / p64.size
        G32.close()
        res32["speedup_vs_fp64"] = res32["cell_updates_per_s"] / r64["cell_updates_per_s"]
    G64.close()

    # the whole configured run (200 steps by default) per precision: band
    # size range, throughput, and the FP32 errors at its end
    full = {}
    if args.full_steps > 0:
        ph64, full["fp64"] = run_full("fp64", args.size, args.grains, args.sigma, args.full_steps)
        if not args.no_fp32:
            ph32, full["fp32"] = run_full("fp32", args.size, args.grains, args.sigma, args.full_steps)
            f32 = full["fp32"]
            f32["label_disagreement_vs_fp64"] = float(np.count_nonzero(ph64 != ph32)) / ph64.size
            f32["speedup_vs_fp64"] = f32["cell_updates_per_s"] / full["fp64"]["cell_updates_per_s"]
            f32["kappa_error_at_end"] = kappa_error_on(ph64, args.size, args.sigma)
    if res32 is not None:
        from gpk_amd import voronoi_host
        phv, _ = voronoi_host(args.size, args.size, args.size, args.grains, 42)
        res32.update(kappa_error_on(phv, args.size, args.sigma))

    cpu = None
    if not args.no_cpu:
        from gpk_amd import voronoi_host
        phi0, g0 = voronoi_host(args.size, args.size, args.size, args.grains, 42)
        cpu = cpu_baseline(phi0, g0, args.size, args.sigma, args.cpu_steps)

    flops = r64["geometry_flops_per_launch"]
    geo_s = r64["geometry_ms_avg"] * 1e-3
    eff = flops / geo_s / 1e12 if geo_s > 0 else 0.0
    # FP64 operations the separable kernel executes per band cell in 3-D:
    # per window row one add per extra chunk plus 6 FMAs, per window plane 9 FMAs
    W = 2 * (int(np.ceil(3.5 * args.sigma)) + 2) + 1
    nch = (W + 8) // 9
    band_per_launch = flops / (2.0 * W ** 3 * 9)
    exec_flops = band_per_launch * (W * W * (3 * (nch - 1) + 12) + W * 18)
    exec_rate = exec_flops / geo_s / 1e12 if geo_s > 0 else 0.0
    traffic = None
    traffic_src = None
    tpath = os.path.join(ROOT, "profiles", "geometry_fp64_256_traffic.json")
    if os.path.exists(tpath) and args.size == 256 and args.grains == 20000:
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
        traffic_src = "%s (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch; %s)" % (
            tj.get("source"), tj.get("measured", "undated"))
    out = {
        "metric": "interface cell-updates/s, 256^3 grain growth: FP32 vs FP64 GPU vs CPU",
        "value": r64["cell_updates_per_s"],
        "unit": "interface cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (periodic Voronoi, %d sites, seed 42)" % args.grains,
        "config": {"workload": "3-D isotropic grain growth, %d^3 periodic grid, %d Voronoi grains, "
                               "sigma %.2f cells, FP64" % (args.size, args.grains, args.sigma),
                   "grid": [args.size] * 3, "grains": args.grains, "sigma_cells": args.sigma,
                   "precision": "fp64", "parallelism": "single GPU"},
        # the dominant kernel executes FP64 vector (VALU) arithmetic; the
        # masked stencil is not a contraction and uses no matrix cores
        "roofline": {"bound": "valu", "compute": "FP64 vector ALU (VALU), no MFMA",
                     "kernel": "k_geometry_tiled<double,3,8> (separable row-mask tables)",
                     "limiter": "LDS table reads (indexed LDS operations 70% of CU-cycles, bank "
                                "conflicts 37%; profiles/pmc_geometry_fp64_256.md)",
                     "achieved": exec_rate, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                     "frac": exec_rate / PEAK_TFLOPS["fp64"], "traffic": traffic,
                     "traffic_source": traffic_src,
                     "flops_basis": "FP64 operations the kernel executes: per band cell %d window rows x "
                                    "(6 FMA + %d adds) + %d planes x 9 FMA" % (W * W, 3 * (nch - 1), W),
                     "executed_flops_per_launch": exec_flops, "avg_launch_ms": r64["geometry_ms_avg"],
                     "effective": {"tflops": eff, "frac": eff / PEAK_TFLOPS["fp64"],
                                   "flops_per_launch": flops,
                                   "basis": "reference dense formulation, 2 x %d^3 x 9 flops per band cell; "
                                            "above 1 because the separable kernel does not execute it" % W}},
        "cpu_baseline": cpu,
        "fp64": r64,
        "fp32": res32,
        "full_run": full or None,
    }
    if cpu:
        out["speedup_vs_cpu"] = r64["cell_updates_per_s"] / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
