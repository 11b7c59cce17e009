#!/usr/bin/env python3
"""Benchmark: RHS-solves/sec on the N=4000 hexagonal-monolayer frequency sweep.

Workload (one "step"): solve all 10 frequencies x 3 polarisations = 30
right-hand sides of the mutual-polarisation system A(omega) p = E0 for a 2-D
hexagonal monolayer of N = 4000 spheres (80 columns x 50 rows, radius 5 nm,
nearest-neighbour spacing 11 nm, Gaussian position jitter sigma = 0.2 nm with
numpy seed 2, box 880 x 476.31 x 250 nm periodic), Drude ITO-like
eps(omega) = 4 - 1.7^2 / (omega^2 + i 0.08 omega), omega = 0.2 ... 0.8 eV,
eps_m = 1, Ewald xi = 0.5 with errortol = tol = 1e-6 (grid 210 x 114 x 61,
9^3 stencil), GMRES(50) with left Jacobi preconditioning, stopping rule of
the reference (||M^-1 (b - A x)|| < 1e-6 ||b||).

value = RHS-solves/s of the batched mixed-precision solve (all 30 systems
in one batched GMRES; FP32 reciprocal-space applies inside the Arnoldi
cycles, FP64 residual b - A x at every restart), measured over the device
solve phase with inputs resident in device memory.  Reported beside it: the
plain per-frequency FP64 solve (reference order, one system at a time,
derivative initial guesses) and its ratio, the accuracy of the batched
dipoles against the FP64 solution, per-kernel-class HIP event timings with
the HBM roofline of the dominant class, and the CPU oracle on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RHS-solves/sec, N=4000 hex monolayer, batched+mixed vs plain FP64 at 1e-6"
HBM_PEAK_GBS = 8000.0


def hex_monolayer_config():
    cols, rows, s, a = 80, 50, 11.0, 5.0
    j, i = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    x = (i + 0.5 * (j % 2)) * s
    y = j * s * np.sqrt(3) / 2
    pos = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], axis=1)
    pos = pos + np.random.default_rng(2).normal(0.0, 0.2, pos.shape)
    box = np.array([cols * s, rows * s * np.sqrt(3) / 2, 250.0])
    omega = np.linspace(0.2, 0.8, 10)
    eps = 4.0 - 1.7 ** 2 / (omega ** 2 + 1j * omega * 0.08)
    return pos, box, a, omega, eps


def drude(omega):
    return 4.0 - 1.7 ** 2 / (omega ** 2 + 1j * omega * 0.08)


def secondary_workload(name):
    """The other benchmark workloads (parity cases; timed with --workload): returns
    (positions, solver keyword arguments, right-hand sides per sweep, description)."""
    if name == "sc512":
        g = np.arange(8) * 3.0 + 1.5
        pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        kw = dict(box=[24.0] * 3, eps_p=-2.0 + 0.5j, radius=1.0, E0=[[1.0, 0.0, 0.0]], tol=1e-8, restart=30,
                  grid=(32, 32, 32), support=8)
        return pos, kw, 1, "simple cubic 8^3, spacing 3, L=24, eps=-2+0.5i, E0=x, grid 32^3, support 8, tol 1e-8"
    if name == "hex4000-quads":
        pos, box, a, omega, eps = hex_monolayer_config()
        kw = dict(box=box, eps_p=eps, radius=a, tol=1e-6, field_type="monodisperse", quadrupoles=True)
        return pos, kw, 30, "hex monolayer N=4000 with quadrupoles on every particle (n=32000), 10 x 3 RHS"
    if name == "poly1000":
        rng = np.random.default_rng(3)
        L = 76.35
        rad = rng.uniform(1.0, 4.0, 1000)
        pos = np.zeros((1000, 3))
        k = 0
        while k < 1000:  # random sequential addition, periodic
            c = rng.uniform(0.0, L, 3)
            d = pos[:k] - c
            d -= L * np.round(d / L)
            if np.all(np.einsum("ij,ij->i", d, d) >= (rad[:k] + rad[k]) ** 2):
                pos[k] = c
                k += 1
        kw = dict(box=[L] * 3, eps_p=drude(np.linspace(0.3, 1.0, 20)), radius=rad, tol=1e-6)
        return pos, kw, 60, "polydisperse N=1000, radii U(1,4) nm, RSA in 76.35 nm cube, 20 x 3 RHS"
    if name == "random100k":
        L, n1 = 111.769, 47
        rng = np.random.default_rng(5)
        g = (np.arange(n1) + 0.5) * (L / n1)
        sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        pos = sites[rng.permutation(len(sites))[:100000]] + rng.uniform(-0.18, 0.18, (100000, 3))
        kw = dict(box=[L] * 3, eps_p=np.full(4, -3.0 + 0.3j), radius=1.0, xi=np.sqrt(-np.log(1e-6)) / 4.0,
                  tol=1e-6, grid=(256, 256, 256), support=8)
        return pos, kw, 12, ("N=100000, a=1, volume fraction 0.3 (jittered lattice subset), 256^3, support 8, "
                             "rc=4, 4 identical frequencies x 3 RHS")
    raise SystemExit(f"unknown workload {name}")


def run_secondary(cumpm_amd, name, steps, solver_type, profile_only=False, autotune=False):
    """Batched refined sweep (timed over `steps` solves after one warm-up) beside the plain
    per-frequency FP64 sweep in reference order; one JSON line.  profile_only: the batched
    refined solves only (the run a profiler wraps)."""
    pos, kw, nrhs, desc = secondary_workload(name)
    out = {}
    from cumpm_amd import _capi
    s = None
    legs = [("batched_refined", "refined", True, steps), ("plain_fp64", "double", False, 1)]
    if profile_only:
        s = cumpm_amd.dipole_solver(quiet=True, precision="refined", batched=True, solver_type=solver_type,
                                    autotune=autotune, **kw)
        for _ in range(steps + 1):
            s.reset()
            s.compute(pos)
        print(json.dumps({"workload": desc, "solves": steps + 1, "ms_per_sweep": s.get_timing()[0]}))
        return
    for label, prec, batched, n in legs:
        s = None  # the previous solver's device memory is released before this one is measured
        base = _capi.memory_stats(reset_peak=True)[0]
        s = cumpm_amd.dipole_solver(quiet=True, precision=prec, batched=batched, solver_type=solver_type,
                                    autotune=autotune, **kw)
        s.compute(pos)
        peak = _capi.memory_stats()[1] - base
        if autotune:
            out[label + "_autotune"] = s.get_tune_info()
        ms = []
        for _ in range(n):
            s.reset()
            s.compute(pos)
            ms.append(s.get_timing()[0])
        reps = [r for w in s.get_reports()[0] for r in w]
        it = [r["iters"] for r in reps]
        out[label] = {"value": nrhs / (np.mean(ms) / 1e3), "ms_per_sweep": float(np.mean(ms)),
                      "iters_min": min(it), "iters_max": max(it),
                      "converged": all(r["converged"] for r in reps),
                      "max_relres_pre": max(r["relres_pre"] for r in reps),
                      "peak_device_bytes": peak}
        out[label + "_dips"] = s.get_dipoles(physical=False)
    dref = out.pop("plain_fp64_dips")
    dmix = out.pop("batched_refined_dips")
    line = {"metric": f"RHS-solves/sec, {name} ({solver_type}), batched+mixed vs plain FP64", "unit": "RHS-solves/s",
            "value": out["batched_refined"]["value"], "higher_is_better": True, "n_gpus": 1, "steps": steps,
            "data": "synthetic", "config": {"workload": desc},
            "speedup_vs_plain_fp64": out["batched_refined"]["value"] / out["plain_fp64"]["value"],
            "dipole_diff_rel_to_max": float(np.abs(dmix - dref).max() / np.abs(dref).max()), **out}
    print(json.dumps(line))


def make_solver(cumpm_amd, box, eps, a, precision, batched, tol, autotune=False):
    return cumpm_amd.dipole_solver(box, eps, radius=a, eps_m=1.0, xi=0.5, tol=tol, quiet=True,
                                   solver_type="gmres", field_type="monodisperse", precision=precision,
                                   batched=batched, autotune=autotune)


def run_once(s, pos):
    s.reset()
    s.compute(pos)
    solve_ms, setup_ms, total_ms = s.get_timing()
    return solve_ms, setup_ms, total_ms


def algorithmic_bytes(cls, info, N, B, fp32, pairs):
    """Algorithmic HBM bytes of one launch of a kernel class over B right-hand sides.

    The reciprocal grid is the full nx*ny*nz grid, or in the slab layout only the
    occupied z-planes (slab_planes * nx * ny voxels).  "fft_axisK" is one pass of
    the transform (one read and one write of the grid); "contract" reads the grid
    once and produces the reciprocal field of every particle (3 complex FP64 per
    particle and RHS); "zline" is the full grid's z transforms and Green's-tensor
    projection in one pass; "zfused" is the slab Green's-tensor pass (grid read and
    write plus its coefficient table); "xtoeplitz" is the same with the slab's x
    transform and inverse x transform fused in (still one read and one write); "gather" reads the dipoles of the
    neighbours and of the particle and writes the result (pair data on the fly
    counts 20 B per pair)."""
    nx, ny, nz = list(info.grid)
    V = nx * ny * nz
    Vp = info.slab_planes * nx * ny if info.slab_planes else V
    c = 8 if fp32 else 16
    r = 4 if fp32 else 8
    W = info.stencil_width
    grid = 3 * Vp * B * c
    if cls == "spread":
        return grid + 48 * N * B + N * (3 * W * r + 12)
    if cls.startswith("fft_axis"):
        return 2 * grid
    if cls == "fft":
        return 2 * 2 * grid if info.slab_planes else 3 * 2 * grid
    if cls == "xtoeplitz":  # x transform + plane coupling + inverse x transform: one read, one write
        return 2 * grid + 6 * info.slab_planes * nx * ny * c
    if cls == "zfused":
        return 2 * grid + (6 * info.slab_planes * nx * ny * c if info.slab_planes else 4 * V * r)
    if cls in ("scale", "zline"):  # zline: z transforms + projection, still one read and one write
        return 2 * grid + 4 * V * r
    if cls == "contract":
        return grid + 48 * N * B + N * (3 * W * r + 12)
    if cls == "gather":
        return 2 * 48 * N * B + 16 * N * B + pairs * 20 + 24 * N
    return 0


def kernel_class(name):
    """Kernel class of a kernel name as profilers print it (template arguments included);
    "other" (no algorithmic byte model) for kernels outside the dipole operator's path."""
    base = name.split("<")[0]
    return {"spread_mfma_kernel": "spread", "spread_kernel": "spread", "line_fft2_kernel": "fft_axis",
            "line_fft_kernel": "fft_axis", "ztoeplitz_kernel": "zfused", "zfused_kernel": "zfused", "xtoeplitz_kernel": "xtoeplitz",
            "scale_kernel": "scale", "contract_mfma_kernel": "contract", "contract_nb_mfma_kernel": "contract",
            "contract_full_mfma_kernel": "contract", "zline_scale_kernel": "zline",
            "finish_kernel": "gather", "gather_kernel": "gather"}.get(base, "other")


def kernel_bytes(name, info, N, B, pairs):
    return algorithmic_bytes(kernel_class(name), info, N, B, "<float" in name, pairs)


def _pmc_run(counter, args, outdir, name, timeout_s=120):
    """One rocprofv3 --pmc pass of a single counter over `python3 bench.py <args>` in a child
    process (this process stays idle meanwhile); returns {kernel name: [value per dispatch]}."""
    import csv
    import subprocess
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", outdir, "-o", name, "--",
           sys.executable, os.path.abspath(__file__)] + args
    proc = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, start_new_session=True)
    try:
        _, err = proc.communicate(timeout=timeout_s)
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, 9)
        proc.communicate()
        raise RuntimeError(f"rocprofv3 --pmc {counter} exceeded {timeout_s} s")
    if proc.returncode != 0:
        raise RuntimeError(f"rocprofv3 --pmc {counter} failed ({proc.returncode}): {err.decode()[-300:]}")
    out = {}
    with open(os.path.join(outdir, f"{name}_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def measure_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (a profiler-style name without the argument list) and
    the right-hand sides per dispatch they were measured at: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) from two rocprofv3 passes over one --profile-only solve of this workload, each
    scaled by what the same counter reads for the library's plain copy kernel of known size
    (--copy-calibration)."""
    import tempfile
    with tempfile.TemporaryDirectory(prefix="bench_pmc_") as d:
        prof = ["--profile-only", "--no-autotune", "--steps", "1", "--warmup", "1"]
        cal = ["--copy-calibration"]
        copy_bytes = float(256 << 20)
        scale = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            vals = [v for k, vs in _pmc_run(c, cal, d, "cal_" + c).items() if "profile_copy_kernel" in k for v in vs]
            scale[c] = copy_bytes / (np.mean(vals) * 1024.0)
        total = 0.0
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            vals = [v for k, vs in _pmc_run(c, prof, d, "op_" + c).items()
                    if k.replace("void mpm::", "").split("(")[0] == kernel for v in vs]
            if not vals:
                raise RuntimeError(f"no {kernel} dispatch in the --pmc {c} pass")
            total += np.mean(vals) * 1024.0 * scale[c]
    return total, scale


def cpu_baseline(pos, box, a, eps, iters_needed, budget_s=20.0):
    """Time the oracle's FP64 GMRES on one right-hand side (omega index 5, x polarisation)
    for a capped number of Arnoldi steps, then scale to a full solve using the number of
    steps the same algorithm needs for that system (from the FP64 GPU solve)."""
    from oracle import oracle as O
    op = O.Operator(box / a, 1e-6, 0.5)
    t0 = time.perf_counter()
    op.set_positions(pos / a)
    t_setup = time.perf_counter() - t0
    w = 5
    sc = O.self_coefficient(np.full(len(pos), eps[w]))
    b = np.zeros((len(pos), 3), complex)
    b[:, 0] = 1.0
    # one apply to size the sample
    t0 = time.perf_counter()
    op.apply(np.zeros((len(pos), 3), complex) + 1.0, sc)
    t_apply = time.perf_counter() - t0
    steps = int(max(2, min(iters_needed, budget_s / max(t_apply, 1e-3) - 2)))
    t0 = time.perf_counter()
    x, rep = op.solve(b, sc, tol=1e-6, method="gmres", restart=50, maxiter=steps)
    t_solve = time.perf_counter() - t0
    per_matvec = t_solve / max(rep["matvecs"] + 1, 1)  # +1: the diagnostic final residual
    full = per_matvec * (iters_needed + 2)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": 1.0 / full,
        "unit": "RHS-solves/s",
        "cores": O.num_threads(),
        "cpu_model": model,
        "nproc": os.cpu_count(),
        "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
        "kind": "port",
        "sample": (f"oracle FP64 GMRES(50)+Jacobi on 1 RHS (omega index {w}, x-pol) capped at {steps} "
                   f"Arnoldi steps ({t_solve:.1f} s, {per_matvec * 1e3:.0f} ms per operator apply incl. "
                   f"orthogonalisation); scaled to the {iters_needed} steps the FP64 solve needs; "
                   f"per-frame set-up {t_setup:.2f} s not included"),
    }


def copy_calibration(_capi, mib=256, reps=20):
    """Launch the library's plain copy kernel on `mib` MiB (read once, written once) and print
    its duration: under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per dispatch, divided by the
    true bytes, is the factor that turns the counters of the operator's kernels into bytes."""
    import ctypes
    lib = _capi.load()
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    _capi.check(lib.mpm_buffer_alloc(nbytes, ctypes.byref(src)))
    _capi.check(lib.mpm_buffer_alloc(nbytes, ctypes.byref(dst)))
    try:
        host = np.ones(nbytes // 8)
        _capi.check(lib.mpm_buffer_upload(src, host.ctypes.data, nbytes))
        del host
        for _ in range(5):
            _capi.check(lib.mpm_profile_copy(dst, src, nbytes, None))
        _capi.check(lib.mpm_device_synchronize())
        t0 = time.perf_counter()
        for _ in range(reps):
            _capi.check(lib.mpm_profile_copy(dst, src, nbytes, None))
        _capi.check(lib.mpm_device_synchronize())
        dt = (time.perf_counter() - t0) / reps
    finally:
        lib.mpm_buffer_free(src)
        lib.mpm_buffer_free(dst)
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": reps + 5,
                      "us_per_launch": dt * 1e6, "GBps": 2 * nbytes / dt / 1e9}))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 reference-order solve")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure roofline.traffic")
    ap.add_argument("--tune-grid", action="store_true",
                    help="secondary workloads: let the autotuner also grow the grid to 7-smooth lengths "
                         "(a finer discretisation, within the Ewald tolerance)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default layout and FFT instead of timing the candidates on the first frame")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + steps of the batched solve only (for rocprofv3)")
    ap.add_argument("--workload", default="hex4000",
                    choices=["hex4000", "sc512", "hex4000-quads", "poly1000", "random100k"],
                    help="hex4000 (default) is the headline measurement; the others time the remaining "
                         "benchmark workloads (batched refined vs plain FP64) and print one JSON line")
    ap.add_argument("--solver", default="gmres", choices=["gmres", "bicgstab"],
                    help="Krylov method of a --workload other than hex4000")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (the known-traffic kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the sweep's 30 independent systems fit one MI355X, and "
              "this library has no multi-GPU path (sharding would split the right-hand sides with no "
              "communication).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import cumpm_amd
    from cumpm_amd import _capi

    if args.copy_calibration:
        copy_calibration(_capi)
        return
    if args.workload != "hex4000":
        run_secondary(cumpm_amd, args.workload, max(args.steps, 1), args.solver, args.profile_only,
                      "grid" if args.tune_grid else (not args.no_autotune))
        return

    pos, box, a, omega, eps = hex_monolayer_config()
    N = len(pos)
    B = len(omega) * 3
    log = lambda *m: print("[bench]", *m, file=sys.stderr, flush=True)

    # ---- batched + mixed (refined) ----
    tune = not args.no_autotune
    s_mix = make_solver(cumpm_amd, box, eps, a, "refined", True, args.tol, tune)
    if args.profile_only:  # per-kernel launch and RHS counts of every solve below
        _capi.profile_reset()
        _capi.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(max(args.warmup, 1)):
        run_once(s_mix, pos)
    log(f"warmup done in {time.perf_counter() - t0:.1f} s")
    _capi.check(_capi.load().mpm_device_synchronize())
    solve_ms = []
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        sm, _, _ = run_once(s_mix, pos)
        solve_ms.append(sm)
    _capi.check(_capi.load().mpm_device_synchronize())
    wall = time.perf_counter() - wall0
    ms_step = float(np.mean(solve_ms))
    value = B / (ms_step / 1e3)
    reps_mix = s_mix.get_reports()[0]
    dips_mix = s_mix.get_dipoles(physical=False)
    info = s_mix.get_op_info()
    log(f"batched+refined: {ms_step:.1f} ms/step -> {value:.1f} RHS-solves/s")
    if args.profile_only:
        _capi.profile_enable(False)
        kern = {}
        for n in _capi.profile_names():
            if "<" in n:
                ms, cnt, units = _capi.profile_get(n)
                kern[n] = {"launches": cnt, "mean_rhs": units / max(cnt, 1)}
        print(json.dumps({"metric": METRIC, "value": value, "ms_per_step": ms_step,
                          "matvecs_per_solve": sum(r["matvecs"] for w in reps_mix for r in w),
                          "solves": max(args.warmup, 1) + args.steps, "kernels": kern}))
        return

    # ---- kernel-class timings (HIP events on the launch stream) ----
    _capi.profile_reset()
    _capi.profile_enable(True)
    run_once(s_mix, pos)
    _capi.profile_enable(False)
    classes, kernels = {}, {}
    for name in _capi.profile_names():
        ms, cnt, units = _capi.profile_get(name)
        if cnt:
            (kernels if "<" in name else classes)[name] = {"total_ms": ms, "launches": cnt,
                                                           "mean_rhs": units / cnt}

    # ---- plain FP64, reference order ----
    plain = None
    if not args.no_plain:
        s_plain = make_solver(cumpm_amd, box, eps, a, "double", False, args.tol, tune)
        run_once(s_plain, pos)  # warm the FFT plans
        p_ms, _, _ = run_once(s_plain, pos)
        reps_plain = s_plain.get_reports()[0]
        dips_plain = s_plain.get_dipoles(physical=False)
        plain = {
            "value": B / (p_ms / 1e3),
            "ms_per_sweep": p_ms,
            "iters_per_rhs": [r["iters"] for w in reps_plain for r in w],
            "max_relres_true": max(r["relres_true"] for w in reps_plain for r in w),
        }
        log(f"plain fp64: {p_ms:.1f} ms/sweep -> {plain['value']:.1f} RHS-solves/s")
        # batched FP64 (same batched code path, FP64 applies only)
        s_b64 = make_solver(cumpm_amd, box, eps, a, "double", True, args.tol, tune)
        run_once(s_b64, pos)
        b64_ms, _, _ = run_once(s_b64, pos)
        plain["batched_fp64_value"] = B / (b64_ms / 1e3)
        dd = dips_mix - dips_plain
        per_rhs = np.linalg.norm(dd.reshape(len(omega), N * 3, 3).transpose(0, 2, 1).reshape(B, -1), axis=1) / \
            np.linalg.norm(dips_plain.reshape(len(omega), N * 3, 3).transpose(0, 2, 1).reshape(B, -1), axis=1)
        accuracy = {
            "max_abs_dipole_diff_rel_to_max": float(np.abs(dd).max() / np.abs(dips_plain).max()),
            "max_per_rhs_rel_l2": float(per_rhs.max()),
            "batched_max_relres_true": max(r["relres_true"] for w in reps_mix for r in w),
            "batched_max_relres_pre": max(r["relres_pre"] for w in reps_mix for r in w),
        }
    else:
        accuracy = None

    # ---- roofline of the dominant kernel (one template instance, HIP events) ----
    pairs = int(info.num_pairs)
    for name, v in kernels.items():
        v["avg_ms"] = v["total_ms"] / v["launches"]
        v["algorithmic_GBps"] = kernel_bytes(name, info, N, v["mean_rhs"], pairs) / (v["avg_ms"] / 1e3) / 1e9
    for c, v in classes.items():
        v["avg_ms"] = v["total_ms"] / v["launches"]
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    kd = kernels[dom]
    bytes_per_launch = kernel_bytes(dom, info, N, kd["mean_rhs"], pairs)
    achieved = bytes_per_launch / (kd["avg_ms"] / 1e3) / 1e9
    traffic, traffic_note = None, None
    if not args.no_traffic:
        # the --profile-only solve is this run's solve (same systems, same deferral), so its
        # dispatches of `dom` carry the same right-hand sides per launch on average
        try:
            traffic, tscale = measure_traffic(dom)
            traffic_note = (f"rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch of {dom} over one --profile-only "
                            f"solve, scaled by the copy-kernel factors {tscale['FETCH_SIZE']:.3f} (fetch) and "
                            f"{tscale['WRITE_SIZE']:.3f} (write)")
            log(f"traffic: {traffic / 1e6:.1f} MB per launch ({traffic / bytes_per_launch:.2f}x algorithmic)")
        except Exception as e:  # no profiler on this machine, or a failed pass: reported, not fatal
            traffic_note = f"not measured: {e}"
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note, "kernel": dom,
                "algorithmic_bytes_per_launch": bytes_per_launch, "avg_launch_ms": kd["avg_ms"],
                "mean_rhs_per_launch": kd["mean_rhs"], "launches": kd["launches"]}

    # ---- CPU oracle ----
    cpu = None
    if not args.no_cpu:
        iters_needed = plain["iters_per_rhs"][5 * 3] if plain else 60
        cpu = cpu_baseline(pos, box, a, eps, iters_needed)
        log(f"cpu oracle: {cpu['value']:.4f} RHS-solves/s ({cpu['cores']} threads)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "RHS-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 reciprocal inside Arnoldi, c128 residual/Krylov",
        "data": "synthetic",
        "config": {
            "workload": "hex monolayer N=4000 (80x50, a=5 nm, spacing 11 nm, jitter 0.2 nm seed 2, "
                        "box 880x476.31x250 nm), Drude ITO eps_inf=4 wp=1.7 eV gamma=0.08 eV, "
                        "10 freqs 0.2-0.8 eV x 3 pol = 30 RHS, GMRES(50)+Jacobi, tol 1e-6",
            "grid": list(info.grid), "stencil": int(info.stencil_width), "rc": info.rc,
            "slab_planes": int(info.slab_planes), "fft": "built-in" if info.fft_backend == 0 else "rocFFT",
            "rhs": B, "particles": N, "parallelism": "single GPU, all RHS batched",
            "autotune": s_mix.get_tune_info() if tune else None,
        },
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (value / plain["value"]) if plain else None,
        "accuracy_vs_fp64": accuracy,
        "batched_iters_per_rhs": [r["iters"] for w in reps_mix for r in w],
        "batched_matvecs_fp64_per_rhs": [r["matvecs_fp64"] for w in reps_mix for r in w],
        "end_to_end_wall_rhs_per_s": B * args.steps / wall,
        "kernel_classes": classes,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
