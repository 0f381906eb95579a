#!/usr/bin/env python
"""Forward + adjoint PM steps per second on one MI355X.

Workload (default): 256^3 particles on a 256^3 mesh, periodic box of
1024 Mpc/h, cosmology Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96,
sigma8=0.8; Gaussian initial conditions from seeded white noise and the
Eisenstein-Hu P(k), 2LPT at a=0.1, then the kick-drift-kick leapfrog to
a=1.0.  One timed "step" is one leapfrog step through the forward pass and
the hand-written adjoint (checkpoints every 5 steps, so the backward sweep
recomputes each segment), with the loss sum((rho - rho_t)^2)/sigma^2 of the
final painted density and its gradient back to the initial positions and
velocities.  Inputs are resident in HBM when the timed region starts.

Prints one JSON line with the throughput, the roofline of the dominant
kernel (timed with HIP events on the launch stream) and a CPU baseline (the
float32 NumPy/SciPy port of the same step timed on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch from hardware counters (FETCH_SIZE and WRITE_SIZE
# passes over `python bench.py --workload traffic`, each scaled by the same
# counter's reading for a plain copy of a known 64 MiB: x1.999 for reads,
# x1.000 for writes), by timed kernel name and mesh side, with the bench's
# sort tuning.  The commands and the counter values are in
# profiles/sorted_gathers_256_traffic.md and kernel_bounds_256_counters.md.
PMC_TRAFFIC_BYTES = {("bin_kick_kernel[drift]", 256): 1038.1e6,
                     ("bin_grad_kernel[drift]", 256): 1468.0e6,
                     ("bin_accum_kernel[nc=1]", 256): 361.8e6,
                     ("bin_accum_kernel[nc=3]", 256): 848.1e6}


def fixed_tuning(args):
    """The sort tuning codes given on the command line, (0, 0) with
    --no-tune, or None: run the autotuner."""
    if args.tuning:
        return tuple(int(v) for v in args.tuning.split(","))
    return (0, 0) if args.no_tune else None


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed leapfrog steps (forward + adjoint)")
    ap.add_argument("--warmup", type=int, default=3, help="untimed warm-up steps")
    ap.add_argument("--mesh", type=int, default=256)
    ap.add_argument("--box", type=float, default=1024.0)
    ap.add_argument("--checkpoint", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle", action="store_true", help="skip the float64-oracle error checks")
    ap.add_argument("--no-batch48", action="store_true",
                    help="skip the 48-simulation batch against the plain one-by-one form")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--workload", choices=["pm", "diffsim", "batch48", "large", "kernels", "traffic"], default="pm",
                    help="pm: the default forward+adjoint step benchmark; diffsim: the differentiable "
                         "128^3 simulation with gradients w.r.t. the noise and (Omega_c, sigma8); batch48: 48 "
                         "batched simulations; large: 512^3 particles on a 1024^3 mesh; kernels: the paint / "
                         "read / pm_forces microbenchmark; traffic: fixed launches for hardware-counter passes")
    ap.add_argument("--no-tune", action="store_true", help="skip the sort autotuner (default widths)")
    ap.add_argument("--tuning", type=str, default=None,
                    help="fixed sort tuning 'chunk,slice,rho_global,xcd_slabs' codes instead of the autotuner")
    ap.add_argument("--plain-steps", type=int, default=5,
                    help="leapfrog steps of the plain stored-trajectory autodiff comparison (0: skip)")
    return ap.parse_args()


def kernel_time_ms(fn, reps):
    """Average duration of fn's launches, HIP events on torch's current stream
    (the stream libjaxpm_hip launches on)."""
    import torch
    fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, x, v, runs=3):
    """One forward + adjoint step of the float32 NumPy/SciPy port on the
    host: one warm-up step, then the median of ``runs`` timed steps."""
    from oracle import cpu_step
    cores = len(os.sched_getaffinity(0))
    shape = (n, n, n)
    r = np.random.default_rng(0)
    cx = r.normal(size=x.shape).astype(np.float32)
    cv = r.normal(size=x.shape).astype(np.float32)
    times = []
    for i in range(runs + 1):
        t0 = time.perf_counter()
        cpu_step.step_fwd_adj(x, v, cx, cv, shape, 0.01, 0.01, cores)
        if i:
            times.append(time.perf_counter() - t0)
        print(f"bench: cpu baseline step {i} ({'timed' if i else 'warm-up'}) {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    dt = float(np.median(times))
    return {"value": 1.0 / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"forward+adjoint leapfrog steps at {n}^3 (drift, force, kick, recomputed force, force VJP, "
                      f"drift adjoint) of oracle/cpu_step.py, float32 NumPy/SciPy c2c FFT with {cores} workers: "
                      f"1 warm-up step, then the median of {runs} steps "
                      f"({', '.join(f'{t:.1f}' for t in times)} s)"}


def config1_forward(dev):
    """The CPU-runnable reference case: 32^3 particles on a 32^3 mesh, box
    256 Mpc/h, 1LPT at a=0.1 and 10 leapfrog steps to a=1, CIC paint of the
    final positions.  Times the GPU forward and the float64 oracle's on the
    same noise (default_rng(0)), and reports the per-cell RMS difference of
    the final densities relative to the oracle's density contrast."""
    import torch

    from jaxpm_amd import integrate, painting, pm
    from jaxpm_amd import cosmology as pc
    from oracle import cosmology as oc
    from oracle import ode as oo
    from oracle import pm as opm
    shape, box = (32, 32, 32), (256.0,) * 3
    cp = dict(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    steps = np.linspace(0.1, 1.0, 11)
    noise = np.random.default_rng(0).standard_normal(shape, dtype=np.float32)
    ktab = np.logspace(-4, 1, 128)

    def gpu_run():
        cosmo = pc.Cosmology(**cp)
        with torch.no_grad():
            ic = pm.linear_field(shape, box, pm.TabulatedPk(ktab, pc.linear_matter_power(cosmo, ktab)),
                                 torch.as_tensor(noise, device=dev))
            lat = torch.stack(torch.meshgrid(*[torch.arange(32, device=dev, dtype=torch.float32)] * 3,
                                             indexing="ij"), -1)
            dx, p, _ = pm.lpt(cosmo, ic, lat, a=0.1, order=1)
            x, _ = integrate.leapfrog((lat + dx).reshape(-1, 3), p.reshape(-1, 3), steps, cosmo, shape)
            return painting.cic_paint(torch.zeros(shape, device=dev), x)

    gpu_run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rho = gpu_run()
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    c = oc.Cosmology(**cp)
    pk = oc.linear_matter_power(c, ktab)
    ic = opm.linear_field(shape, box, lambda k: np.interp(k, ktab, pk), noise.astype(np.float64))
    parts = opm.uniform_particles(shape).astype(np.float64)
    dx, p, _ = opm.lpt(c, ic, parts, a=0.1, order=1)
    x, _ = oo.leapfrog((parts + dx).reshape(-1, 3), p.reshape(-1, 3), steps, c, shape, reuse_force=True)
    ref = oo.final_density(x, shape)
    t_cpu = time.perf_counter() - t0
    r = rho.double().cpu().numpy()
    return {"workload": "32^3 particles on 32^3 mesh, box 256 Mpc/h, 1LPT a=0.1 + 10 leapfrog steps + CIC paint",
            "gpu_s": round(t_gpu, 4), "cpu_oracle_fp64_s": round(t_cpu, 3),
            "density_rms_rel_diff_vs_fp64": float(np.sqrt(np.mean((r - ref) ** 2)) / np.sqrt(np.mean((ref - 1) ** 2))),
            "pk_max_rel_diff_vs_fp64": pk_rel_diff(rho, ref, box)}


def pk_rel_diff(rho_gpu, rho_ref, box):
    """Largest relative difference over the default k bins between the
    GPU's power spectrum of a density (jaxpm_amd.utils, on the device) and
    the float64 oracle's of the oracle's density (oracle.utils, NumPy)."""
    from jaxpm_amd import utils as pu
    from oracle import utils as ou
    _, pg = pu.power_spectrum(rho_gpu, box_shape=box)
    _, pr = ou.power_spectrum(rho_ref, box_shape=box)
    pg = np.asarray(pg.cpu() if hasattr(pg, "cpu") else pg, np.float64)
    return float(np.max(np.abs(pg - pr) / np.abs(pr)))


def diffsim(args):
    """Differentiable single simulation: 128^3 particles and mesh, box
    512 Mpc/h, 2LPT at a=0.1 + 40 leapfrog steps to a=1, checkpoints every 8
    steps; loss sum((delta - delta_t)^2)/0.1^2 against the final density of
    seed-4 noise; gradients w.r.t. the seed-3 white noise and (Omega_c,
    sigma8).  One timed unit is one forward + gradient simulation."""
    import torch

    from jaxpm_amd import _lib, diffsim
    from jaxpm_amd import cosmology as pc
    dev = _lib.device()
    n = 128 if args.mesh == 256 else args.mesh
    shape = (n, n, n)
    S = 40
    model = diffsim.LptNBody(shape, (4.0 * n,) * 3, np.linspace(0.1, 1.0, S + 1),
                             base_cosmo=pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8),
                             order=2, checkpoint_every=8)
    noise = torch.as_tensor(np.random.default_rng(3).standard_normal(shape, dtype=np.float32), device=dev)
    with torch.no_grad():
        tnoise = torch.as_tensor(np.random.default_rng(4).standard_normal(shape, dtype=np.float32), device=dev)
        target = model(tnoise, model.parameter_vector()) - 1.0

    def run():
        nt = noise.clone().requires_grad_(True)
        params = model.parameter_vector().requires_grad_(True)
        rho = model(nt, params)
        loss = ((rho - 1.0 - target) ** 2).sum() / 0.1 ** 2
        loss.backward()
        return float(loss.detach()), nt.grad, params.grad

    for _ in range(max(args.warmup, 1) if args.warmup > 0 else 0):
        run()
    torch.cuda.synchronize()
    reps = max(1, min(args.steps, 5))
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, gn, gp = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    out = {"metric": "forward+gradient simulations/s, 128^3 particles and mesh, 2LPT + 40 leapfrog steps",
           "value": round(1.0 / dt, 4), "unit": "sims/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
           "ms_per_step": round(1000 * dt, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"differentiable simulation {n}^3, box {4 * n} Mpc/h, 2LPT a=0.1, {S} steps, "
                                  "checkpoint every 8, gradients w.r.t. noise and (Omega_c, sigma8)"},
           "loss": loss, "grad_noise_rms": float(gn.pow(2).mean().sqrt()),
           "grad_params": {"Omega_c": float(gp[0]), "sigma8": float(gp[1])},
           "peak_mem_gb": round(torch.cuda.max_memory_allocated() / 1e9, 2)}
    print(json.dumps(out))


def batch48(args):
    """The batch48 workload's JSON line (batch48_result)."""
    print(json.dumps(batch48_result(args)))


def batch48_result(args):
    """Batched forward+adjoint: 48 simulations, 16 each on 64^3, 96^3 and
    128^3 meshes with one particle per cell, box 4 n Mpc/h, Omega_c ~
    U[0.2, 0.35] and sigma8 ~ U[0.7, 0.9] (default_rng(6)), 1LPT at a=0.1 and
    20 leapfrog steps to a=1, loss sum_b sum((delta_b - delta_t,b)^2)/0.1^2,
    gradients w.r.t. every white-noise field.  The N-body stage runs as one
    launch per stage over the batch and replays captured HIP graphs; it is
    timed against the same 48 simulations run one by one, with the custom
    adjoint and in the plain stored-trajectory form."""
    import torch

    from jaxpm_amd import _lib, diffsim
    from jaxpm_amd import cosmology as pc
    dev = _lib.device()
    sizes = [64] * 16 + [96] * 16 + [128] * 16
    r = np.random.default_rng(6)
    om = r.uniform(0.2, 0.35, len(sizes))
    s8 = r.uniform(0.7, 0.9, len(sizes))
    cosmos = [pc.Cosmology(Omega_c=float(a), Omega_b=0.05, h=0.7, n_s=0.96, sigma8=float(b)) for a, b in zip(om, s8)]
    shapes = [(n, n, n) for n in sizes]
    boxes = [(4.0 * n,) * 3 for n in sizes]
    a_steps = np.linspace(0.1, 1.0, 21)
    nr = np.random.default_rng(7)
    noises = [torch.as_tensor(nr.standard_normal(sh, dtype=np.float32), device=dev) for sh in shapes]
    tnoise = [torch.as_tensor(nr.standard_normal(sh, dtype=np.float32), device=dev) for sh in shapes]
    tuning = fixed_tuning(args)
    if tuning is None:
        from jaxpm_amd import autotune
        lat = torch.cat([torch.stack(torch.meshgrid(*[torch.arange(n, device=dev, dtype=torch.float32)] * 3,
                                                     indexing="ij"), -1).reshape(-1, 3) for n in sizes])
        tuning, _ = autotune.tune_batch(shapes, [n ** 3 for n in sizes], lat + 0.3)
        del lat
    t0 = time.perf_counter()
    model = diffsim.BatchLptNBody(shapes, boxes, a_steps, cosmos, order=1, checkpoint_every=5, graph=True,
                                  sort_tuning=tuning)
    t_build = time.perf_counter() - t0
    noise_cat = torch.cat([nz.reshape(-1) for nz in noises])
    with torch.no_grad():
        t_cat = model.forward_fixed_cosmology(torch.cat([t.reshape(-1) for t in tnoise])) - 1.0
    cells = [n ** 3 for n in sizes]
    offs = np.concatenate([[0], np.cumsum(cells)])
    targets = [t_cat[offs[b]:offs[b + 1]].reshape(shapes[b]) for b in range(len(sizes))]

    def batched():
        nt = noise_cat.clone().requires_grad_(True)
        rho = model.forward_fixed_cosmology(nt)
        loss = ((rho - 1.0 - t_cat) ** 2).sum() / 0.1 ** 2
        loss.backward()
        return [nt.grad[offs[b]:offs[b + 1]].reshape(shapes[b]) for b in range(len(sizes))]

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps, out

    reps = max(1, min(args.steps, 5))
    t_batch, g_batch = timed(batched, reps)
    print(f"batch48: batched forward+adjoint {t_batch:.3f} s per pass", file=sys.stderr, flush=True)
    out = {"metric": "forward+adjoint simulations/s, 48 simulations (16 each at 64^3, 96^3, 128^3), 1LPT + 20 "
                     "leapfrog steps", "value": round(len(sizes) / t_batch, 3), "unit": "sims/s", "n_gpus": 1,
           "steps": reps, "warmup": 1, "ms_per_step": round(1000 * t_batch, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batched forward+adjoint of 48 simulations, 16 each at 64^3/96^3/128^3, box 4n "
                                  "Mpc/h, Omega_c~U[0.2,0.35], sigma8~U[0.7,0.9], 1LPT a=0.1 + 20 steps, "
                                  "checkpoint every 5; every stage batched, the N-body sweeps graph-captured",
                      "particles": int(sum(n ** 3 for n in sizes))},
           "graph_capture_s": round(t_build, 2), "sort_tuning": list(tuning)}
    if args.plain_steps > 0:
        for method in ("custom", "plain"):
            singles = [diffsim.LptNBody(sh, bx, a_steps, base_cosmo=c, order=1, checkpoint_every=5, method=method)
                       for sh, bx, c in zip(shapes, boxes, cosmos)]

            def one_by_one():
                grads = []
                for m, nz, t in zip(singles, noises, targets):
                    nt = nz.clone().requires_grad_(True)
                    rho = m(nt, m.parameter_vector())
                    (((rho - 1.0 - t) ** 2).sum() / 0.1 ** 2).backward()
                    grads.append(nt.grad)
                return grads

            t_one, g_one = timed(one_by_one, 1)
            print(f"batch48: one by one ({method}) {t_one:.3f} s", file=sys.stderr, flush=True)
            rel = max(float((a - b).abs().max() / b.abs().max()) for a, b in zip(g_batch, g_one))
            rms = max(float((a - b).square().mean().sqrt() / b.square().mean().sqrt()) for a, b in zip(g_batch, g_one))
            out[f"one_by_one_{method}"] = {"sims_per_s": round(len(sizes) / t_one, 3),
                                           "speedup_of_batched": round(t_one / t_batch, 3),
                                           "grad_max_rel_diff": rel, "grad_rms_rel_diff": rms}
            del singles
            torch.cuda.empty_cache()
    out["peak_mem_gb"] = round(torch.cuda.max_memory_allocated() / 1e9, 2)
    del model
    torch.cuda.empty_cache()
    return out


def kernels(args):
    """Kernel microbenchmark, forward only: cic_paint, cic_read, one
    pm_forces and one lens-plane density (1024^2, with and without smoothing)
    on a 256^3 mesh with 256^3 particles, in two position states --
    uniform random in the box, and clustered (2LPT at a=0.1 from seed-2 noise,
    box 1024 Mpc/h, then 20 leapfrog steps to a=1).  GB/s on algorithmic
    bytes: paint 12N + 4M (positions in, mesh out), read 12N + 4M + 4N.
    Checks that need no oracle (parity with the oracle is the test suite's):
    the paint conserves mass, the binned paint equals the atomic one, and
    pm_forces of the tile-sorted pipeline equals the unfused one (atomic
    paint, three force meshes, separate read)."""
    import torch

    from jaxpm_amd import _lib, _ops, integrate, lensing, painting, pm
    from jaxpm_amd import cosmology as pc
    dev = _lib.device()
    n = args.mesh
    shape = (n, n, n)
    N = M = n ** 3
    gen = torch.Generator(device=dev).manual_seed(2)
    uniform = torch.rand((N, 3), generator=gen, device=dev) * n
    cosmo = pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    k, pk = pc.pk_table(cosmo)
    with torch.no_grad():
        ic = pm.linear_field(shape, (args.box,) * 3, pm.TabulatedPk(k, pk), torch.randn(shape, generator=gen,
                                                                                          device=dev))
        lattice = torch.stack(torch.meshgrid(*[torch.arange(n, device=dev, dtype=torch.float32)] * 3,
                                             indexing="ij"), -1)
        dx, p, _ = pm.lpt(cosmo, ic, lattice, a=0.1, order=2)
        x0 = (lattice + dx).reshape(-1, 3).contiguous()
        v0 = p.reshape(-1, 3).contiguous()
        st = integrate.Stepper(shape, N)
        sched = integrate.schedule_tensor(integrate.kdk_schedule(cosmo, np.linspace(0.1, 1.0, 21)))
        clustered, _ = integrate.NBody.apply(x0, v0, sched, st, 5)
    del ic, lattice, dx, p, x0, v0, st
    reps = args.kernel_reps
    paint_bytes, read_bytes = 12 * N + 4 * M, 12 * N + 4 * M + 4 * N
    plane_bytes = 12 * N + 4 * 1024 ** 2           # positions in, plane out
    out_states = {}
    for name, pos in (("uniform", uniform), ("clustered", clustered)):
        mesh = torch.zeros(shape, device=dev)
        t_paint = kernel_time_ms(lambda: painting.cic_paint(mesh, pos), reps)
        t_atomic = kernel_time_ms(lambda: _ops.paint(torch.zeros(shape, device=dev), pos, method="atomic",
                                                     zero=True), reps)
        rho = painting.cic_paint(mesh, pos)
        rho_atomic = _ops.paint(torch.zeros(shape, device=dev), pos, method="atomic", zero=True)
        t_read = kernel_time_ms(lambda: painting.cic_read(rho, pos), reps)
        t_read_sorted = kernel_time_ms(lambda: painting.cic_read(rho, pos, method="sorted"), reps)
        same_read = bool(torch.equal(painting.cic_read(rho, pos), painting.cic_read(rho, pos, method="sorted")))
        t_forces = kernel_time_ms(lambda: pm.pm_forces(pos, mesh_shape=shape), reps)
        f_fused = pm.pm_forces(pos, mesh_shape=shape)
        f_plain = _ops.pm_forces_fused(pos, shape, atomic=True)
        # lens plane of the central quarter of the box at 1024^2 (lensing.py:11-44)
        R, center, width = 1024, n / 2, n / 4
        t_plane = kernel_time_ms(lambda: lensing.density_plane(pos, shape, center, width, R), reps)
        t_plane_s = kernel_time_ms(lambda: lensing.density_plane(pos, shape, center, width, R, smoothing_sigma=1.0),
                                   reps)
        plane = lensing.density_plane(pos, shape, center, width, R)
        in_slab = int(((pos[:, 2] > center - width / 2) & (pos[:, 2] <= center + width / 2)).sum())
        plane_mass = float(plane.double().sum()) * (n / R) ** 2 * width
        torch.cuda.synchronize()
        out_states[name] = {
            "cic_paint": {"ms": round(t_paint, 4), "GB_s": round(paint_bytes / (t_paint * 1e-3) / 1e9, 1),
                          "frac_hbm": round(paint_bytes / (t_paint * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "what": "tile sort + LDS windows + face reduce, deterministic"},
            "cic_paint_atomic": {"ms": round(t_atomic, 4),
                                 "GB_s": round(paint_bytes / (t_atomic * 1e-3) / 1e9, 1),
                                 "frac_hbm": round(paint_bytes / (t_atomic * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "what": "one global float atomic per corner"},
            "cic_read": {"ms": round(t_read, 4), "GB_s": round(read_bytes / (t_read * 1e-3) / 1e9, 1),
                         "frac_hbm": round(read_bytes / (t_read * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "what": "direct 8-corner gather"},
            "cic_read_sorted": {"ms": round(t_read_sorted, 4),
                                "GB_s": round(read_bytes / (t_read_sorted * 1e-3) / 1e9, 1),
                                "frac_hbm": round(read_bytes / (t_read_sorted * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                "what": "tile sort + LDS windows (method='sorted')"},
            "pm_forces": {"ms": round(t_forces, 4)},
            "density_plane": {"ms": round(t_plane, 4), "GB_s": round(plane_bytes / (t_plane * 1e-3) / 1e9, 1),
                              "frac_hbm": round(plane_bytes / (t_plane * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                              "what": f"{R}^2 plane of the slab {center - width / 2:g} < z <= "
                                      f"{center + width / 2:g}: wrap, select, 2-D CIC paint, normalise"},
            "density_plane_smoothed": {"ms": round(t_plane_s, 4), "what": "the same plus gaussian_smoothing, "
                                       "sigma 1 pixel (2-D R2C, filter, C2R)"},
            "checks": {"mass_rel_err": abs(float(rho.double().sum()) - N) / N,
                       "density_plane_mass_rel_err": abs(plane_mass - in_slab) / max(in_slab, 1),
                       "sorted_read_bitwise_equal": same_read,
                       "binned_vs_atomic_max_rel": float((rho - rho_atomic).abs().max() / rho.abs().max()),
                       "pm_forces_sorted_vs_unfused_max_rel": float((f_fused - f_plain).abs().max() /
                                                                    f_plain.abs().max())}}
    c = out_states["clustered"]
    out = {"metric": "CIC paint / read GB/s (as % HBM peak) and pm_forces ms at 256^3, uniform and clustered",
           "value": c["cic_paint"]["GB_s"], "unit": "GB/s", "n_gpus": 1, "steps": reps, "warmup": 1,
           "ms_per_step": c["cic_paint"]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"cic_paint / cic_read / pm_forces, {n}^3 particles on {n}^3 mesh, uniform "
                                  "random and 2LPT + 20-step clustered positions", "mesh": n, "particles": N},
           "states": out_states}
    print(json.dumps(out))


def large(args):
    """Large single-GPU run: 512^3 particles (a lattice of spacing 2 cells)
    on a 1024^3 mesh, box 1000 Mpc/h, 2LPT at a=0.1 and 30 leapfrog steps to
    a=1, checkpoints every 5 steps, forward + adjoint with the gradient of
    sum(delta^2)/0.1^2 w.r.t. the 1024^3 white-noise field.  Reports steps/s
    and peak device memory (torch's peak allocation and the device's used
    memory, which includes the hipFFT plans' work areas)."""
    import torch

    from jaxpm_amd import _autograd, _lib, diffsim, integrate
    from jaxpm_amd import cosmology as pc
    dev = _lib.device()
    nm, npart, S = 1024, 512, 30
    shape = (nm, nm, nm)
    cosmo = pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    spec = diffsim.SimSpec(shape, (1000.0,) * 3, np.linspace(0.1, 1.0, S + 1), base_cosmo=cosmo, order=2)
    gen = torch.Generator(device=dev).manual_seed(5)
    noise = torch.randn(shape, generator=gen, device=dev)
    ax = torch.arange(npart, device=dev, dtype=torch.float32) * 2.0
    lattice = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1)
    ktab = torch.as_tensor(spec.ktab, dtype=torch.float32, device=dev)
    stepper = integrate.Stepper(shape, npart ** 3)
    params = spec.parameter_vector()
    mean = npart ** 3 / nm ** 3

    def run():
        nt = noise.clone().requires_grad_(True)
        q, x0, v0 = spec.initial_state(nt, params, True, lattice, ktab)
        xS, _ = integrate.NBody.apply(x0, v0, q[spec.slices["kdk"]], stepper, 5)
        rho = _autograd.CicPaint.apply(torch.zeros(shape, device=dev), xS, None, 1.0, False, False)
        loss = ((rho / mean - 1.0) ** 2).sum() / 0.1 ** 2
        loss.backward()
        return float(loss.detach()), nt.grad

    t0 = time.perf_counter()
    run()                      # plans, tables
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    print(f"large: first forward+adjoint run (with plan set-up) {t_first:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    loss, g = run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    free, total = torch.cuda.mem_get_info()
    out = {"metric": "fwd+adjoint leapfrog steps/s, 512^3 particles on a 1024^3 mesh (incl. 2LPT and its adjoint)",
           "value": round(S / dt, 3), "unit": "steps/s", "n_gpus": 1, "steps": S, "warmup": 1,
           "ms_per_step": round(1000 * dt / S, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "512^3 particles on 1024^3 mesh, box 1000 Mpc/h, 2LPT a=0.1 + 30 steps, checkpoint "
                                  "every 5, gradient w.r.t. the 1024^3 noise", "particles": npart ** 3, "mesh": nm},
           "seconds_per_simulation": round(dt, 3), "first_run_s": round(t_first, 2), "loss": loss,
           "grad_finite": bool(torch.isfinite(g).all()),
           "peak_torch_gb": round(torch.cuda.max_memory_allocated() / 1e9, 2),
           "device_used_gb_after": round((total - free) / 1e9, 2), "device_total_gb": round(total / 1e9, 1)}
    print(json.dumps(out))


def traffic(args):
    """A fixed set of launches for hardware-counter passes (rocprofv3 --pmc
    FETCH_SIZE, then WRITE_SIZE): a plain device copy of 4 n^3 bytes (the
    calibration launch), then on the n^3 workload the tile-sorted force
    evaluations of a 10-step leapfrog from 2LPT initial conditions to the
    clustered state at a=1, on the particles in the tile order of the
    initial positions as NBody runs them, each kick carrying the next drift
    (the last one a plain kick), and 5 force VJPs with the drift adjoint
    there: the launches of the leapfrog's forward and adjoint sweeps.  The
    per-dispatch counters of a kernel divided by those of the copy give its
    HBM traffic in units of the copy's known byte count."""
    import torch

    from jaxpm_amd import _lib, integrate, pm
    from jaxpm_amd import cosmology as pc
    n = args.mesh
    tuning = fixed_tuning(args) or (0, 0)
    dev = _lib.device()
    shape = (n, n, n)
    N = n ** 3
    a = torch.randn(N, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    cosmo = pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    k, pk = pc.pk_table(cosmo)
    gen = torch.Generator(device=dev).manual_seed(2)
    with torch.no_grad():
        ic = pm.linear_field(shape, (4.0 * n,) * 3, pm.TabulatedPk(k, pk), torch.randn(shape, generator=gen,
                                                                                      device=dev))
        lattice = torch.stack(torch.meshgrid(*[torch.arange(n, device=dev, dtype=torch.float32)] * 3,
                                             indexing="ij"), -1)
        dx, p, _ = pm.lpt(cosmo, ic, lattice, a=0.1, order=2)
        x = (lattice + dx).reshape(-1, 3).contiguous()
        v = p.reshape(-1, 3).contiguous()
        st = integrate.Stepper(shape, N, sort_tuning=tuning)
        if st.reorder:
            x, v = st.permute(st.tile_order(x), x, v)
        k0, beta, gamma = integrate.kdk_schedule(cosmo, np.linspace(0.1, 1.0, 11))
        field = st.new_field(keep_sort=True)
        S = len(beta)
        st.field_kick_drift(x, v, k0, beta[0], field, x)
        for i in range(S):
            if i + 1 < S:
                st.field_kick_drift(x, v, gamma[i], beta[i + 1], field, x)
            else:
                st.field_kick(x, v, gamma[i], field)
        cot = torch.randn((N, 3), generator=gen, device=dev)
        gpos = torch.zeros_like(x)
        for _ in range(5):
            st.vjp_field_drift(field, x, cot, float(gamma[-1]), float(beta[-1]), gpos)
    torch.cuda.synchronize()
    print(f"traffic launches done: copy of {4 * N} bytes, {S + 1} force evaluations and 5 force VJPs at {n}^3, "
          f"sort tuning {tuning}", flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the reference's sharded "
              "halo exchange and pencil FFTs are out of scope); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "diffsim":
        return diffsim(args)
    if args.workload == "batch48":
        return batch48(args)
    if args.workload == "large":
        return large(args)
    if args.workload == "kernels":
        return kernels(args)
    if args.workload == "traffic":
        return traffic(args)
    import torch

    from jaxpm_amd import _lib, _ops, integrate, painting, pm
    from jaxpm_amd import cosmology as pc

    dev = _lib.device()
    n = args.mesh
    shape = (n, n, n)
    N = n ** 3
    box = (args.box,) * 3
    cosmo = pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    k, pk = pc.pk_table(cosmo)
    gen = torch.Generator(device=dev).manual_seed(2)
    noise = torch.randn(shape, generator=gen, device=dev)
    with torch.no_grad():
        ic = pm.linear_field(shape, box, pm.TabulatedPk(k, pk), noise)
        lattice = torch.stack(torch.meshgrid(*[torch.arange(n, device=dev, dtype=torch.float32)] * 3,
                                             indexing="ij"), -1)
        dx, p, _ = pm.lpt(cosmo, ic, lattice, a=0.1, order=2)
        x0 = (lattice + dx).reshape(-1, 3).contiguous()
        v0 = p.reshape(-1, 3).contiguous()
        target = painting.cic_paint(torch.zeros(shape, device=dev),
                                    lattice.reshape(-1, 3) + 0.5 * torch.randn((N, 3), generator=gen, device=dev))
    sigma = 0.1

    tuning = fixed_tuning(args)
    if tuning is None:
        from jaxpm_amd import autotune
        tuning, _ = autotune.tune_stepper(shape, x0)
    stepper = integrate.Stepper(shape, N, sort_tuning=tuning)

    def fwd_adj(S, trace=None, x_init=None):
        steps = np.linspace(0.1, 1.0, S + 1)
        sched = integrate.schedule_tensor(integrate.kdk_schedule(cosmo, steps))
        stepper.trace = trace
        xt = (x0 if x_init is None else x_init).clone().requires_grad_(True)
        vt = v0.clone().requires_grad_(True)
        x, v = integrate.NBody.apply(xt, vt, sched, stepper, args.checkpoint)
        rho = painting.cic_paint(torch.zeros(shape, device=dev), x)
        loss = ((rho - target) ** 2).sum() / sigma ** 2
        loss.backward()
        stepper.trace = None
        return x.detach(), xt.grad

    if args.warmup > 0:
        fwd_adj(args.warmup)
    # the timed region: K steps with the library's per-kernel timing off
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xS, g0 = fwd_adj(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert torch.isfinite(g0).all(), "non-finite gradient"
    steps_per_s = args.steps / elapsed
    print(f"bench: timed region {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    # the same K steps again with per-kernel timing on: HIP events recorded by
    # the library on the launch stream around each of its kernels give the
    # average launch durations (the events cost host time per launch, so this
    # pass is not the one `value` comes from)
    trace = {}
    _lib.timing_report()          # drop anything recorded so far
    _lib.timing(True)
    t1 = time.perf_counter()
    fwd_adj(args.steps, trace)
    torch.cuda.synchronize()
    elapsed_timed = time.perf_counter() - t1
    _lib.timing(False)
    per_kernel = _lib.timing_report()
    M = N
    H = n * n * (n // 2 + 1)
    win = 4 * M * (17 / 16) ** 3   # tile windows (16^3 tiles + 1-cell CIC halo)
    faces = 4 * (M // 16 ** 3) * (2 * 721 + 1156)
    algo = {
        "bin_kick_kernel": 36 * N + 4 * M,             # pos, vel r+w, potential
        "bin_kick_kernel[drift]": 48 * N + 4 * M,      # ... and the drifted positions
        "bin_kick_kernel[forces]": 24 * N + 4 * M,     # pos, forces, potential
        "bin_grad_kernel": 48 * N + 8 * M,             # pos, cot, gpos r+w, potential, rho_bar
        "bin_grad_kernel[drift]": 60 * N + 8 * M,      # ... and the drift adjoint's cot write
        "bin_count_kernel": 12 * N,
        "bin_scatter_kernel": 12 * N + 16 * N,
        "bin_accum_kernel[nc=1]": 16 * N + win,
        "bin_accum_kernel[nc=3]": 16 * N + 12 * N + 3 * win,
        # the face reduce of single-chunk tiles (every tile here): per 16^3
        # tile and channel its 721 low-face cells read and written, and the
        # 1156 halo cells of its window slot read (by the upper neighbours)
        "bin_reduce_kernel[nc=1]": faces,
        "bin_reduce_kernel[nc=3]": 3 * faces,
        "potential_kernel": 16 * H,
        "fd4_divergence_kernel": 16 * M,
        "axpy_kernel": 36 * N,
        "hipfft_r2c": 4 * M + 8 * H,
        "hipfft_c2r": 8 * H + 4 * M,
        # the library's FFT, pass by pass
        "fft_r2c_rows_kernel": 4 * M + 8 * H,          # reals in, half spectrum out
        "fft_r2c_rows_kernel[div]": 12 * M + 8 * H,    # three cotangent meshes in (divergence), half spectrum out
        "fft_cols_kernel": 16 * H,                     # in place
        "fft_cols_potential_kernel": 16 * H,           # in place
        "fft_c2r_rows_kernel": 8 * H + 4 * M,
        # aggregates of the passes above (not kernels; never the roofline kernel)
        "potential_solve (5 passes)": 8 * M + 64 * H,
        "potential_solve[div] (5 passes)": 16 * M + 64 * H,
        "native_r2c (3 passes)": 4 * M + 40 * H,
        "native_c2r (3 passes)": 40 * H + 4 * M,
        "abs_max_kernel": 12 * N,
        "permute_rows_kernel": 2 * 24 * N,             # two (N, 3) arrays, read and written
        "record_order_kernel": 16 * N + 4 * N,
    }
    kinfo = {}
    groups = {}
    for name, (cnt, tot) in per_kernel.items():
        avg = tot / cnt
        entry = {"ms": round(avg, 4), "launches_per_step": round(cnt / args.steps, 2),
                 "ms_per_step": round(tot / args.steps, 4)}
        if name in algo:
            gbs = algo[name] / (avg * 1e-3) / 1e9
            entry.update({"GB_s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)})
            if " (" not in name:
                g = groups.setdefault(name.split("[")[0], {"launches": 0, "ms": 0.0, "bytes": 0.0, "variants": {}})
                g["launches"] += cnt
                g["ms"] += tot
                g["bytes"] += cnt * algo[name]
                g["variants"][name] = cnt
        kinfo[name] = entry
    # the roofline kernel: the kernel function (its variants together) with
    # the most GPU time in the timed region
    dom = max(groups, key=lambda k: groups[k]["ms"])
    gd = groups[dom]
    t_dom = gd["ms"] / gd["launches"]
    bytes_dom = gd["bytes"] / gd["launches"]
    achieved = bytes_dom / (t_dom * 1e-3) / 1e9
    pmc_traffic = None
    meas = [PMC_TRAFFIC_BYTES.get((v, n)) for v in gd["variants"]]
    if all(m is not None for m in meas):
        pmc_traffic = sum(m * c for m, c in zip(meas, gd["variants"].values())) / gd["launches"]
    kernel_groups = {k: {"launches_per_step": round(g["launches"] / args.steps, 2),
                         "ms_per_step": round(g["ms"] / args.steps, 4),
                         "ms_per_launch": round(g["ms"] / g["launches"], 4),
                         "GB_s": round(g["bytes"] / (g["ms"] * 1e-3) / 1e9, 1),
                         "frac_hbm": round(g["bytes"] / (g["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                     for k, g in sorted(groups.items(), key=lambda kv: -kv[1]["ms"])}
    # cic_paint and cic_read alone (the public API, forward), on uniform
    # random positions and on the clustered final state (2LPT + the timed
    # steps), with their float64-oracle errors
    reps = args.kernel_reps
    paint_bytes, read_bytes = 12 * N + 4 * M, 12 * N + 4 * M + 4 * N
    uniform = torch.rand((N, 3), generator=gen, device=dev) * n
    paint_read = {}
    for name, pos in (("uniform", uniform), ("clustered", xS)):
        mesh = torch.zeros(shape, device=dev)
        t_p = kernel_time_ms(lambda: painting.cic_paint(mesh, pos), reps)
        rho = painting.cic_paint(mesh, pos)
        t_r = kernel_time_ms(lambda: painting.cic_read(rho, pos), reps)
        paint_read[name] = {
            "cic_paint": {"ms": round(t_p, 4), "GB_s": round(paint_bytes / (t_p * 1e-3) / 1e9, 1),
                          "frac_hbm": round(paint_bytes / (t_p * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "cic_read": {"ms": round(t_r, 4), "GB_s": round(read_bytes / (t_r * 1e-3) / 1e9, 1),
                         "frac_hbm": round(read_bytes / (t_r * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
        if not args.no_oracle:
            from oracle import painting as op
            p64 = pos.double().cpu().numpy()
            ref = op.cic_paint(np.zeros(shape), p64)
            got = rho.double().cpu().numpy()
            paint_read[name]["cic_paint"]["fp64_oracle_max_rel_err"] = float(np.abs(got - ref).max() /
                                                                            np.abs(ref).max())
            paint_read[name]["cic_paint"]["fp64_oracle_rms_rel_err"] = float(
                np.sqrt(np.mean((got - ref) ** 2)) / np.sqrt(np.mean((ref - ref.mean()) ** 2)))
            paint_read[name]["cic_paint"]["fp64_oracle_pk_max_rel_err"] = pk_rel_diff(rho, ref, (args.box,) * 3)
            rd = painting.cic_read(rho, pos).double().cpu().numpy()
            refr = op.cic_read(got, p64)
            paint_read[name]["cic_read"]["fp64_oracle_max_rel_err"] = float(np.abs(rd - refr).max() /
                                                                           np.abs(refr).max())
            del ref, got, rd, refr, p64
    del uniform
    print("bench: paint / read microbenchmark done", file=sys.stderr, flush=True)
    fp64 = None
    if not args.no_oracle:
        from oracle import pm as opm
        f = pm.pm_forces(xS, mesh_shape=shape).double().cpu().numpy()
        ref = opm.pm_forces(xS.double().cpu().numpy(), mesh_shape=shape)
        fp64 = {"pm_forces_max_rel_err": float(np.abs(f - ref).max() / np.abs(ref).max()),
                "gradient": "the adjoint is checked against float64 central differences of the oracle in "
                            "tests/test_gpu_diffsim.py and tests/test_gpu_configs.py"}
        del f, ref
        fp64["config1"] = config1_forward(dev)
        print("bench: float64 oracle checks done", file=sys.stderr, flush=True)

    # the plain per-simulation form: the same leapfrog as differentiable
    # tensor expressions, autograd over the stored trajectory
    plain = None
    if args.plain_steps > 0:
        P = args.plain_steps

        def plain_fwd_adj():
            sched = integrate.schedule_tensor(integrate.kdk_schedule(cosmo, np.linspace(0.1, 1.0, P + 1)))
            xt = x0.clone().requires_grad_(True)
            vt = v0.clone().requires_grad_(True)
            x, v = integrate.leapfrog_autodiff(xt, vt, sched, shape, trajectory=traj)
            rho = painting.cic_paint(torch.zeros(shape, device=dev), x)
            (((rho - target) ** 2).sum() / sigma ** 2).backward()
            return x.detach(), xt.grad

        traj = None
        plain_fwd_adj()
        torch.cuda.synchronize()
        base_mem = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        t0 = time.perf_counter()
        xp, gp = plain_fwd_adj()
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        mem_plain = torch.cuda.max_memory_allocated() - base_mem
        torch.cuda.reset_peak_memory_stats()
        xc, gc = fwd_adj(P)
        torch.cuda.synchronize()
        mem_custom = torch.cuda.max_memory_allocated() - base_mem
        rel = float((gp - gc).abs().max() / gc.abs().max())
        rms = float((gp - gc).pow(2).mean().sqrt() / gc.pow(2).mean().sqrt())
        # where the two gradients differ: the plain form rerun, keeping the
        # positions of its force evaluations (outside the timed runs)
        traj = []
        plain_fwd_adj()
        gap = integrate.gradient_gap(xp, xc, gp, gc, trajectory=traj)
        traj = None
        # the custom adjoint against itself, with every initial coordinate
        # moved by one float32 ulp: the float32 sensitivity of this gradient
        xn, gn = fwd_adj(P, x_init=torch.nextafter(x0, torch.full_like(x0, float("inf"))))
        gap["custom_vs_custom_1ulp"] = {
            "max_rel": float((gn - gc).abs().max() / gc.abs().max()),
            "rms_rel": float((gn - gc).pow(2).mean().sqrt() / gc.pow(2).mean().sqrt()),
            "final_position_max_diff_cells": float((xn - xc).abs().max())}
        del xn, gn
        plain = {"steps_per_s": round(P / t_plain, 3), "steps": P,
                 "speedup_of_custom_adjoint": round(steps_per_s / (P / t_plain), 3),
                 "grad_max_rel_diff": rel, "grad_rms_rel_diff": rms, "grad_gap": gap,
                 "extra_mem_gb_plain": round(mem_plain / 1e9, 2), "extra_mem_gb_custom": round(mem_custom / 1e9, 2),
                 "what": "same leapfrog as tensor expressions over pm_forces (per-op adjoints), autograd over "
                         "the stored trajectory, no checkpointing; grad w.r.t. x0 compared with the custom adjoint"}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, x0.cpu().numpy(), v0.cpu().numpy())

    # the batched form of the target: the 48-simulation batch against the
    # same 48 simulations one by one in the plain stored-trajectory form
    batched = None
    if not args.no_batch48:
        torch.cuda.empty_cache()
        b = batch48_result(argparse.Namespace(steps=3, tuning=None, no_tune=False, plain_steps=1))
        batched = {"workload": b["config"]["workload"], "sims_per_s": b["value"], "s_per_pass": b["ms_per_step"] / 1e3,
                   "one_by_one_plain": b.get("one_by_one_plain"), "one_by_one_custom": b.get("one_by_one_custom"),
                   "sort_tuning": b["sort_tuning"], "peak_mem_gb": b["peak_mem_gb"]}
        print("bench: batch48 done", file=sys.stderr, flush=True)

    out = {
        "metric": "fwd+adjoint PM steps/s at 256^3 mesh; CIC paint/read GB/s as % HBM peak",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * elapsed / args.steps, 3),
        "ms_per_step_with_kernel_timing": round(1000 * elapsed_timed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"leapfrog forward+adjoint, {n}^3 particles on {n}^3 mesh, box {args.box:g} Mpc/h, "
                               f"2LPT IC at a=0.1, {args.steps} steps to a=1, checkpoint every {args.checkpoint}",
                   "mesh": n, "particles": N, "checkpoint_every": args.checkpoint,
                   "sort_tuning": {"chunk": 2 ** (9 + tuning[0]) if tuning[0] else 8192,
                                   "slice": 2 ** (11 + tuning[1]) if tuning[1] else 8192,
                                   "vjp_density_window": "caches" if len(tuning) > 2 and tuning[2] else "lds",
                                   "xcd_placement": "slabs" if len(tuning) > 3 and tuning[3] else
                                   "round-robin (accumulation), runs of 32 chunks (gathers)"}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if pmc_traffic is None else round(pmc_traffic),
                     "algorithmic_bytes": int(bytes_dom), "ms_per_launch": round(t_dom, 4),
                     "launches_per_step": round(gd["launches"] / args.steps, 2),
                     "variants": {k: round(v / args.steps, 2) for k, v in gd["variants"].items()}},
        "kernel_functions": kernel_groups,
        "kernels": kinfo,
        "paint_read_256": paint_read,
        "fp64_oracle": fp64,
        "plain_autodiff": plain,
        "batched_48": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
