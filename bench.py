"""Benchmark of the FEM elasticity hot path on one MI355X.

Headline (``value``): Mtet-evals/s of pass A -- energy + gradient + Hessian
diagonal of stable Neo-Hookean over a 64^3 x 5 tet cube (1,310,720 tets,
274,625 vertices, FP32, per-tet mu lognormal(median 1e4, sigma 0.5),
lambda = 10 mu), one step = one full pass including the deterministic
vertex reduction and the FP64 energy sum, inputs resident in HBM.

Also reported on the same line: the dominant kernel's roofline (algorithmic
bytes per launch / measured kernel time vs 8 TB/s), the CPU baseline (the C
restatement of the reference's separate fun/grad/hess_diag passes, timed on
the host), PNCG solves/s at 1e-6 relative gradient norm, the ARAP pass, and a
batched-vs-plain solve comparison.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mtet-evals/s (energy+grad+Hdiag) & PNCG solves/s, batched+adjoint vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_cube_problem(n, material, dtype, device, seed_mu=1, seed_noise=0, noise_frac=0.1, chunks=1, seed_act=8):
    """Config-2 style cube: lognormal mu, lambda = 10 mu, gravity, z = 0 pinned, noisy start
    (muscle: per-tet activation, normal with standard deviation 0.1 per component)."""
    import torch

    from apple_amd import fem
    from apple_amd.mesh import Region, cube

    m = cube(n)
    rng = np.random.default_rng(seed_mu)
    mu = 1e4 * np.exp(0.5 * rng.standard_normal(m.n_cells))
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10.0 * mu
    region = Region.from_mesh(m)
    if material == "muscle":
        region.cell_data["activation"] = 0.1 * np.random.default_rng(seed_act).standard_normal((m.n_cells, 6))
    cls = {"snh": fem.StableNeoHookean, "arap": fem.Arap, "muscle": fem.StableNeoHookeanMuscle}[material]
    pot = cls.from_region(region, dtype=dtype, device=device, chunks=chunks)
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8 * 1e3 * region.dV / 4.0, 4))
    fixed = np.zeros((m.n_points, 3), bool)
    fixed[m.points[:, 2] == 0.0] = True
    x0 = noise_frac / n * np.random.default_rng(seed_noise).standard_normal((m.n_points, 3))
    x0[fixed] = 0.0
    t = lambda a: torch.as_tensor(a, dtype=dtype, device=device).contiguous()
    return dict(mesh=m, region=region, pot=pot, f=t(f), fixed=fixed, x0=t(x0), mu=mu)


class PassBench:
    """One element pass of a potential, prepared for timing: outputs
    allocated, materials bound and every pointer converted up front, so that
    ``run`` is a single call into the library (``apl_bench_elem_pass``:
    ``iters`` back-to-back launches between HIP events on the stream)."""

    def __init__(self, pot, u, what):
        import torch

        from apple_amd import _lib
        from apple_amd.runtime import current_stream, ptr

        m = pot.mesh
        self._keep = a = torch.empty(m.n_verts, 3, dtype=m.dtype, device=m.device)
        self._keep_b = b = torch.empty_like(a)
        self._keep_e = e = torch.zeros(m.n_inst, dtype=torch.float64, device=m.device)  # one energy per instance
        lam = pot.materials.get("lambda")
        getattr(pot, "_prepare", lambda: None)()  # binds per-mesh material arrays (the muscle activation)
        p = u if what != _lib.APL_ENERGY | _lib.APL_GRAD | _lib.APL_HESS_DIAG else None
        self._fn = _lib.lib().apl_bench_elem_pass
        self._graph_fn = _lib.lib().apl_bench_elem_graph
        self._args = (m._handle, int(pot.material), what, ptr(pot.materials["mu"]),
                      ptr(lam) if lam is not None else None, ptr(u), ptr(p), ptr(e), ptr(a), ptr(b))
        self._stream = current_stream(m.device)
        self._k = ctypes.c_double()
        self._tp = ctypes.c_double()
        self._k_ref = ctypes.byref(self._k)
        self._tp_ref = ctypes.byref(self._tp)
        self._check = _lib.check

    def run(self, iters, kernel=True, full=True):
        """(mean element-kernel ms, mean full-pass ms), None where skipped."""
        self._check(self._fn(*self._args, int(iters), self._k_ref if kernel else None,
                             self._tp_ref if full else None, self._stream))
        return (self._k.value if kernel else None), (self._tp.value if full else None)

    def run_graph(self, iters):
        """Mean whole-pass ms of ``iters`` passes replayed from one HIP graph
        (``apl_bench_elem_graph``); the first call with a given ``iters``
        captures the graph, so call once untimed before timing."""
        self._check(self._graph_fn(*self._args, int(iters), self._tp_ref, self._stream))
        return self._tp.value


def bench_pass(pot, u, what, iters, kernel=True, full=True):
    """(mean element-kernel ms, mean full-pass ms) over `iters` back-to-back
    launches each, timed with HIP events on the stream (None when skipped)."""
    return PassBench(pot, u, what).run(iters, kernel, full)


def bench_pcie_inclusive(pot, u, iters=20, warmup=3):
    """Pass A for a caller whose arrays live in host memory: u uploaded from
    pinned memory, the pass writing (not adding into) its outputs, gradient
    and Hessian diagonal downloaded, all on one stream; wall time per pass
    over ``iters`` passes.  Not the headline: the C ABI takes device
    pointers, as the reference's potentials take device arrays."""
    import torch

    T, V = pot.mesh.n_tets, pot.mesh.n_verts
    u_host = torch.empty(V, 3, dtype=u.dtype, pin_memory=True)
    u_host.copy_(u.cpu())
    g_host = torch.empty(V, 3, dtype=u.dtype, pin_memory=True)
    h_host = torch.empty_like(g_host, pin_memory=True)
    u_dev = torch.empty_like(u)
    e = torch.zeros(1, dtype=torch.float64, device=u.device)
    g = torch.zeros_like(u)
    h = torch.zeros_like(u)

    def one():
        u_dev.copy_(u_host, non_blocking=True)
        pot.fused(u_dev, energy=e, grad=g, hess_diag=h, accumulate=False)
        g_host.copy_(g, non_blocking=True)
        h_host.copy_(h, non_blocking=True)

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        one()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / iters * 1e3
    nbytes = 3 * V * u.element_size() * 3
    log(f"[bench] pass A from host memory: {ms * 1e3:.1f} us per pass ({T / (ms / 1e3) / 1e6:.0f} Mtet-evals/s, "
        f"{nbytes / 1e6:.1f} MB over PCIe)")
    return {"mtet_evals_per_s": round(T / (ms / 1e3) / 1e6, 1), "ms_per_pass": round(ms, 5),
            "pcie_bytes_per_pass": nbytes}


def cpu_baseline(prob, target_s=12.0, threads=1):
    """C restatement of the reference passes, timed on the host (bounded sample)."""
    from oracle import c_oracle

    region = prob["region"]
    m = prob["mesh"]
    cp = c_oracle.SnhProblem(m.cells, region.dhdX, region.dV, prob["mu"], 10.0 * prob["mu"])
    used = c_oracle.set_threads(threads)
    u = np.ascontiguousarray(prob["x0"].double().cpu().numpy())
    g = np.zeros_like(u)
    h = np.zeros_like(u)
    cp.pass_a(u, g, h)  # warm
    n = 0
    t0 = time.perf_counter()
    while True:
        cp.pass_a(u, g, h)
        n += 1
        if time.perf_counter() - t0 >= target_s:
            break
    dt = time.perf_counter() - t0
    return dict(
        value=round(n * m.n_cells / dt / 1e6, 4),
        unit="Mtet-evals/s",
        cores=int(used),
        kind="port",
        sample=f"{n} passes of fun+grad+hess_diag (separate passes, atomic scatter, FP64) over the same "
        f"{m.n_cells}-tet mesh, {dt:.1f} s, {used} thread(s)",
        cpu=host_cpu(),
    )


def host_cpus() -> int:
    """Cores this process may run on (the affinity mask, else the CPU count)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def host_cpu() -> dict:
    """The host CPU: model name and the cores available to this process."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "cores_available": host_cpus(), "os_cpu_count": os.cpu_count()}


def load_traffic(name, key="pass_bytes_per_launch"):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE and WRITE_SIZE
    passes, scaled by the copy-kernel calibration); the summary lives under
    profiles/ as text with lines '<key>: <int>'."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.strip().startswith(key + ":"):
                return int(line.split(":", 1)[1].strip())
    return None


def bench_solve(prob, material, dev, rtol=1e-6):
    """One timed PNCG solve of a cube problem after a short graph warm-up."""
    import torch

    from apple_amd.solver import DeviceSolver, PncgConfig, mask_bits

    solver = DeviceSolver(prob["pot"])
    mask = mask_bits(prob["fixed"], dev)
    xw = prob["x0"].clone()
    solver.solve(xw, prob["f"], mask, PncgConfig(rtol=rtol, max_steps=16))
    x = prob["x0"].clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solver.solve(x, prob["f"], mask, PncgConfig(rtol=rtol, max_steps=60000))[0]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"[bench] PNCG {material}: {st.result.name} in {st.n_steps} steps, {dt:.3f} s")
    return {
        "solves_per_s": round(1.0 / dt, 3),
        "solve_s": round(dt, 4),
        "result": st.result.name,
        "steps": st.n_steps,
        "evals": st.n_evals,
        "relative_grad_norm": st.relative_grad_norm,
        "us_per_iteration": round(dt / max(st.n_evals, 1) * 1e6, 2),
        # FP32 element kernels; the solver keeps its own FP64 copy of the
        # iterate (24 B per vertex staged per block slot instead of 12)
        "iterate": "f64 copy" if prob["pot"].mesh.dtype == torch.float32 else "f64",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="cube resolution (cells per side)")
    ap.add_argument("--quick", action="store_true", help="pass timings only (no solves, no CPU baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--chunks", type=int, default=1, help="tet chunks of 256 per element block")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the headline mesh's schedule too (it gains little on one large cube; the batched "
                         "comparison always autotunes its batch)")
    ap.add_argument("--batched-only", action="store_true", help="only the batched forward+adjoint comparison")
    ap.add_argument("--diff-only", action="store_true", help="only the 32^3 adjoint-vs-finite-difference check")
    ap.add_argument("--combined-only", action="store_true", help="only the FP32-vs-FP64 batched forward+adjoint")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--fd-tets", type=int, default=64,
                    help="central-difference tets of the differentiation check in the full run (--diff-only: 64)")
    ap.add_argument("--no-combined", action="store_true")
    ap.add_argument("--no-inverse", action="store_true")
    ap.add_argument("--config1-only", action="store_true", help="only the 10^3 cube solved end to end, CPU and GPU")
    ap.add_argument("--inverse-only", action="store_true", help="only the muscle-activation inverse loop")
    ap.add_argument("--setup-only", action="store_true", help="only the mesh setup timings (device vs host)")
    ap.add_argument("--batched-meshes", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the multi-threaded CPU baseline (0: every core available to the process)")
    ap.add_argument("--profile-pass", action="store_true",
                    help="only run pass-A launches and a reference copy (for rocprofv3 counter runs)")
    ap.add_argument("--profile-hess-prod", action="store_true",
                    help="only 20 pass-C (hess_prod, SNH) launches and 3 calibration copies, for rocprofv3")
    ap.add_argument("--profile-batched", action="store_true",
                    help="only 20 pass-A launches over the 256-cube batch (n ~ U{6..32}) and 3 calibration copies")
    ap.add_argument("--profile-material", choices=("snh", "arap", "muscle"), default="snh",
                    help="material of the --profile-solve run")
    ap.add_argument("--profile-solve", type=int, default=0, metavar="STEPS",
                    help="only run a PNCG solve of STEPS iterations (for rocprofv3 kernel traces)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "apple_amd runs on one GPU: --gpus must be 1 (the element passes would shard by tet slabs with a one-layer vertex halo; not implemented)"}))
        sys.exit(2)

    import torch

    from apple_amd import _lib

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    dtype = torch.float32
    n = args.n
    if args.batched_only:
        print(json.dumps(bench_batched(dev, dtype, n_meshes=args.batched_meshes)), flush=True)
        return
    if args.diff_only:
        print(json.dumps(bench_differentiation(dev)), flush=True)
        return
    if args.combined_only:
        print(json.dumps(bench_combined(dev)), flush=True)
        return
    if args.inverse_only:
        print(json.dumps(bench_inverse(dev)), flush=True)
        return
    if args.setup_only:
        print(json.dumps(bench_setup(dev, args.n)), flush=True)
        return
    if args.cpu_threads <= 0:
        args.cpu_threads = host_cpus()
    if args.config1_only:
        print(json.dumps(bench_config1(dev, args.cpu_threads)), flush=True)
        return
    if args.profile_batched:
        from apple_amd import batch

        insts, _ = make_cube_instances(args.batched_meshes, 6, 32, 2, 3)
        b = batch.Batch(insts, _lib.APL_SNH, dtype=dtype, device=dev)
        pot = b.potential
        u = b.x0
        bench_pass(pot, u, _lib.APL_ENERGY | _lib.APL_GRAD | _lib.APL_HESS_DIAG, 2)  # first-call set-up
        k, pa = bench_pass(pot, u, _lib.APL_ENERGY | _lib.APL_GRAD | _lib.APL_HESS_DIAG, 20)
        T, V = b.n_tets, b.n_verts
        log(f"[bench] batched pass A: {T} tets, {V} verts, kernel {k * 1e3:.1f} us, pass {pa * 1e3:.1f} us, "
            f"{(64 * T + 36 * V) / (pa / 1e3) / 1e9 / HBM_PEAK_GBS:.1%} of HBM (whole pass)")
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)  # 256 MiB copies: counter calibration
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        log("[bench] profile-batched done")
        return
    log(f"[bench] building {n}^3 x 5 cube (SNH, FP32)")
    prob = make_cube_problem(n, "snh", dtype, dev, chunks=args.chunks)
    pot = prob["pot"]
    sched = None
    if args.tune:
        from apple_amd import autotune

        r = prob["region"]
        sched = autotune.tune(r.cells, r.dminv, r.dV, r.mesh.n_points, material=_lib.APL_SNH,
                              mu=pot.materials["mu"], lam=pot.materials["lambda"], dtype=dtype, device=dev)
        log(f"[bench] autotuned schedule: chunks {sched.chunks}, grid {sched.grid_pct}% ({sched.table})")
        if sched.chunks != args.chunks:
            prob = make_cube_problem(n, "snh", dtype, dev, chunks=sched.chunks)
            pot = prob["pot"]
        pot.mesh.set_schedule(sched.grid_pct)
    T, V = pot.mesh.n_tets, pot.mesh.n_verts
    info = pot.mesh.info()
    log(f"[bench] partition: {info['n_blocks']} blocks, nloc_max {info['nloc_max']}, {info['n_partials']} partials, "
        f"{info['n_shared_verts']} shared verts of {V}, grid {info['elem_grid_snh']}")
    what_a = _lib.APL_ENERGY | _lib.APL_GRAD | _lib.APL_HESS_DIAG
    u = prob["x0"]

    if args.profile_pass:
        bench_pass(pot, u, what_a, 20)
        aprob = make_cube_problem(n, "arap", dtype, dev)  # ARAP pass A: its VALU count against the ALU roof
        bench_pass(aprob["pot"], aprob["x0"], what_a, 20)
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)  # 256 MiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] profile-pass done")
        return
    if args.profile_hess_prod:
        bench_pass(pot, u, _lib.APL_HESS_PROD, 20)
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)  # 256 MiB copies: counter calibration
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] profile-hess-prod done")
        return
    if args.profile_solve:
        from apple_amd.solver import DeviceSolver, PncgConfig, mask_bits

        if args.profile_material != "snh":
            prob = make_cube_problem(n, args.profile_material, dtype, dev, chunks=args.chunks)
            pot = prob["pot"]
        solver = DeviceSolver(pot)
        mask = mask_bits(prob["fixed"], dev)
        x = prob["x0"].clone()
        solver.solve(x, prob["f"], mask, PncgConfig(rtol=1e-6, max_steps=16))
        x = prob["x0"].clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve(x, prob["f"], mask, PncgConfig(rtol=1e-6, max_steps=args.profile_solve))[0]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"[bench] profile-solve: {st.n_evals} evaluations in {dt * 1e3:.1f} ms")
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)  # 256 MiB copies: counter calibration
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        return

    # ---------------------------------------------------------------- pass A (headline)
    log(f"[bench] warmup {args.warmup}, timed {args.steps} steps of pass A")
    # Outputs, materials and pointers are set up outside the timed region, and
    # the K timed passes are replayed from one HIP graph captured (and run
    # once) beforehand, as the solver replays its iterations: K eager launch
    # pairs leave ~40 queued commands whose retirement makes the closing
    # torch.cuda.synchronize() cost ~30 us instead of ~5 us.
    pass_a = PassBench(pot, u, what_a)
    # the element kernel alone and the passes launched one by one, means of
    # 300 launches each (~20 ms of work, over which the GPU's clocks settle)
    ms_kernel, ms_pass_eager = pass_a.run(300)
    pass_a.run(max(args.warmup, 1))  # the W warmup steps
    pass_a.run_graph(args.steps)  # capture of the K passes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_pass = pass_a.run_graph(args.steps)  # the timed steps: K whole passes, HIP events around them
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = T / (ms_per_step / 1e3) / 1e6
    # algorithmic bytes of the whole pass (u in, element record in, gradient
    # and Hdiag out) over the whole pass's time (element kernel + vertex sums),
    # both measured here with HIP events on the launch stream
    bytes_a = 64 * T + 36 * V
    achieved = bytes_a / (ms_pass / 1e3) / 1e9
    # the element kernel alone: its compulsory bytes are the element record
    # and u (its partial sums are the reduction's own traffic)
    bytes_elem = 64 * T + 12 * V
    traffic = load_traffic("pass_a_snh64_traffic.md") if n == 64 else None
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": ("profiles/pass_a_snh64_traffic.md: rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of the "
                           "element kernel and the vertex-sum kernel of this workload, scaled by a copy kernel"
                           if traffic is not None else None),
        "kernel": "elem_kernel<SNH, E|G|H, float> + finish_kernel<float, 6> (the whole pass)",
        "pass_ms": round(ms_pass, 5),
        "pass_ms_eager": round(ms_pass_eager, 5),
        "kernel_ms": round(ms_kernel, 5),
        "timing": "the K timed passes replayed from one HIP graph, HIP events around the replay on its stream; "
                  "pass_ms_eager and kernel_ms: means of 300 passes / element kernels launched one by one",
        "algorithmic_bytes_per_launch": bytes_a,
        "element_kernel_only": {"algorithmic_bytes": bytes_elem,
                                "frac": round(bytes_elem / (ms_kernel / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    log(f"[bench] pass A: kernel {ms_kernel*1e3:.1f} us, pass {ms_pass*1e3:.1f} us, wall/step {ms_per_step*1e3:.1f} us,"
        f" {value:.0f} Mtet-evals/s, {achieved:.0f} GB/s ({achieved/HBM_PEAK_GBS:.1%} of HBM)")
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mtet-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"pass A (energy+grad+Hdiag incl. vertex sums), stable Neo-Hookean, {n}^3 cube x 5 tets, "
            "mu lognormal(1e4, 0.5), lambda = 10 mu, seeded noisy displacement",
            "n_tets": T,
            "n_verts": V,
            "n_blocks": info["n_blocks"],
            "nloc_max": info["nloc_max"],
            "n_partials": info["n_partials"],
            "schedule": {"chunks": sched.chunks, "grid_pct": sched.grid_pct} if sched else "default",
            "parallelism": "single GPU",
        },
        "roofline": roofline,
    }

    # ---------------------------------------------------------------- other passes
    k_hq, p_hq = bench_pass(pot, u, _lib.APL_HESS_QUAD, 20)
    out["pass_b_hess_quad"] = {
        "kernel_ms": round(k_hq, 5),
        "hbm_frac": round((64 * T + 24 * V) / (k_hq / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    k_hp, p_hp = bench_pass(pot, u, _lib.APL_HESS_PROD, 20)
    out["pass_c_hess_prod"] = {
        "kernel_ms": round(k_hp, 5),
        "hbm_frac": round((64 * T + 36 * V) / (k_hp / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    if not args.quick:
        out["pcie_inclusive"] = bench_pcie_inclusive(pot, u)
    aprob = None
    if not args.quick:
        log("[bench] ARAP pass A")
        aprob = make_cube_problem(n, "arap", dtype, dev)
        ka, pa = bench_pass(aprob["pot"], aprob["x0"], what_a, 20)
        out["arap_pass_a"] = {
            "mtet_evals_per_s": round(T / (pa / 1e3) / 1e6, 1),
            "kernel_ms": round(ka, 5),
            "hbm_frac": round((60 * T + 36 * V) / (ka / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        }

    mprob = None
    if not args.quick:
        log("[bench] SNH muscle pass A")
        mprob = make_cube_problem(n, "muscle", dtype, dev)
        km, pm = bench_pass(mprob["pot"], mprob["x0"], what_a, 20)
        out["muscle_pass_a"] = {
            "mtet_evals_per_s": round(T / (pm / 1e3) / 1e6, 1),
            "kernel_ms": round(km, 5),
            "hbm_frac": round((88 * T + 36 * V) / (km / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        }

    # ---------------------------------------------------------------- PNCG solves (config 2: SNH, ARAP, muscle)
    if not (args.quick or args.no_solve):
        out["pncg_snh"] = bench_solve(prob, "snh", dev)
        out["pncg_arap"] = bench_solve(aprob, "arap", dev)
        out["pncg_muscle"] = bench_solve(mprob, "muscle", dev)
    del aprob

    # ---------------------------------------------------------------- batched (+adjoint) vs plain
    if not (args.quick or args.no_batch):
        out["batched"] = bench_batched(dev, dtype, n_meshes=args.batched_meshes)

    # ---------------------------------------------------------------- differentiation, FP32 vs FP64
    if not (args.quick or args.no_diff):
        out["differentiation"] = bench_differentiation(dev, n_fd=args.fd_tets)
    if not (args.quick or args.no_combined):
        out["combined"] = bench_combined(dev)
    if not (args.quick or args.no_inverse):
        out["inverse"] = bench_inverse(dev)
    if not args.quick:
        out["setup"] = bench_setup(dev, n)

    # ---------------------------------------------------------------- CPU baseline
    if not (args.quick or args.no_cpu):
        out["config1_solve"] = bench_config1(dev, args.cpu_threads)
        log(f"[bench] CPU baseline (C restatement, {args.cpu_threads} threads and 1 thread, {host_cpu()})")
        out["cpu_baseline"] = cpu_baseline(prob, target_s=12.0, threads=args.cpu_threads)
        out["cpu_baseline_1thread"] = cpu_baseline(prob, target_s=12.0, threads=1)
    print(json.dumps(out), flush=True)


def make_cube_instances(n_meshes, lo, hi, seed_sizes, seed_mat, noise_frac=0.1):
    """Cubes with n ~ U{lo..hi} cells per side (seed_sizes); per-instance mu ~
    U[1e3, 1e5], nu ~ U[0.3, 0.45], lambda = 2 mu nu / (1 - 2 nu) (seed_mat);
    gravity (rho = 1e3, lumped), z = 0 face pinned, start = noise_frac * h *
    N(0, 1) on the free DOFs (seed 100 + k)."""
    from apple_amd import batch
    from apple_amd.mesh import Region, cube

    sizes = np.random.default_rng(seed_sizes).integers(lo, hi + 1, n_meshes)
    mrng = np.random.default_rng(seed_mat)
    cache = {}
    insts = []
    for k, n in enumerate(sizes):
        n = int(n)
        if n not in cache:
            m = cube(n)
            region = Region.from_mesh(m)
            f = np.zeros((m.n_points, 3))
            np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * region.dV / 4, 4))
            fixed = np.zeros((m.n_points, 3), bool)
            fixed[m.points[:, 2] == 0] = True
            cache[n] = (m, region, f, fixed)
        m, region, f, fixed = cache[n]
        mu_i = mrng.uniform(1e3, 1e5)
        nu = mrng.uniform(0.3, 0.45)
        mu = np.full(m.n_cells, mu_i)
        lam = 2 * mu * nu / (1 - 2 * nu)
        x0 = noise_frac / n * np.random.default_rng(100 + k).standard_normal((m.n_points, 3))
        x0[fixed] = 0
        insts.append(batch.Instance(region, mu, lam, f, fixed, x0))
    return insts, [int(n) for n in sizes]


def forward_adjoint(b, cfg, adj_rtol, adjoint=True):
    """Forward PNCG of every instance of batch b, then (adjoint) the PCG solve
    H p = -(x - x0) and the per-tet d L / d(mu, lambda) for
    L = 1/2 ||x - x0||^2.  Returns (x, forward stats, adjoint stats, dmu, dlam)."""
    import torch

    solver = b.solver()
    x = b.x0.clone()
    st = solver.solve(x, b.f_ext, b.mask, cfg)
    if not adjoint:
        return x, st, None, None, None
    rhs = (b.x0 - x).contiguous()
    p = torch.zeros_like(x)
    cs = solver.cg(x, rhs, p, b.mask, rtol=adj_rtol, max_iters=20000)
    dmu = torch.empty(b.n_tets, dtype=b.dtype, device=b.device)
    dlam = torch.empty_like(dmu)
    b.mesh.eval(b.material, b.mu, b.lam, x, p, dmu=dmu, dlam=dlam)
    return x, st, cs, dmu, dlam


def bench_batched(dev, dtype, n_meshes=256, lo=6, hi=32, seed_sizes=2, seed_mat=3, adj_rtol=1e-6):
    """Forward + adjoint of uneven meshes as one batch vs the same work as a
    sequential loop of per-instance GPU solves (setup excluded from both)."""
    import torch

    from apple_amd import _lib, batch
    from apple_amd.solver import PncgConfig

    t_setup = time.perf_counter()
    insts, sizes = make_cube_instances(n_meshes, lo, hi, seed_sizes, seed_mat)
    cfg = PncgConfig(rtol=1e-6, max_steps=20000)
    b = batch.Batch(insts, _lib.APL_SNH, dtype=dtype, device=dev, schedule="auto")
    singles = [batch.Batch([inst], _lib.APL_SNH, dtype=dtype, device=dev) for inst in insts]
    warm = PncgConfig(rtol=1e-6, max_steps=16)
    for bb in [b] + singles:
        forward_adjoint(bb, warm, adj_rtol, adjoint=False)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    res = {}
    for adjoint in (False, True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, stb, csb, _, _ = forward_adjoint(b, cfg, adj_rtol, adjoint)
        torch.cuda.synchronize()
        t_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        ok_fwd, ok_adj = 0, 0
        for bb in singles:
            _, st, cs, _, _ = forward_adjoint(bb, cfg, adj_rtol, adjoint)
            ok_fwd += st[0].success
            ok_adj += cs[0].success if cs is not None else 0
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        key = "forward_adjoint" if adjoint else "forward"
        ok_b = int(sum(s.success for s in stb))
        res[key] = {
            # solves/s count only the instances that reach rtol 1e-6
            "batched_solves_per_s": round(ok_b / t_batch, 3),
            "plain_solves_per_s": round(ok_fwd / t_plain, 3),
            "batched_instances_per_s": round(n_meshes / t_batch, 3),
            "plain_instances_per_s": round(n_meshes / t_plain, 3),
            "speedup": round(t_plain / t_batch, 3),
            "speedup_converged_solves": round((ok_b / t_batch) / max(ok_fwd / t_plain, 1e-300), 3),
            "batched_s": round(t_batch, 4),
            "plain_s": round(t_plain, 4),
            "batched_forward_converged": int(sum(s.success for s in stb)),
            "plain_forward_converged": int(ok_fwd),
        }
        if adjoint:
            res[key]["batched_adjoint_converged"] = int(sum(s.success for s in csb))
            res[key]["plain_adjoint_converged"] = int(ok_adj)
        log(f"[bench] batched {key}: {t_batch:.3f} s vs plain {t_plain:.3f} s ({t_plain / t_batch:.2f}x)")
    res["workload"] = (f"{n_meshes} cubes, n ~ U{{{lo}..{hi}}} (seed {seed_sizes}), {b.n_tets} tets total, SNH FP32, "
                       f"mu ~ U[1e3,1e5], nu ~ U[0.3,0.45] (seed {seed_mat}), PNCG to 1e-6 (STAGNATION after "
                       f"{cfg.stall_steps} steps without progress near it); adjoint: PCG to "
                       f"{adj_rtol:g} on H p = -(x - x0), then the mixed pass for dL/d(mu, lambda); both paths "
                       "replay captured graphs (PNCG and PCG)")
    its = np.array([s.n_steps for s in stb])
    res["batched_iterations"] = {"median": int(np.median(its)), "p95": int(np.percentile(its, 95)),
                                 "max": int(its.max())}
    res["max_n"] = max(sizes)
    res["setup_s"] = round(t_setup, 2)
    res["batch_schedule"] = {"chunks": b.schedule.chunks, "grid_pct": b.schedule.grid_pct,
                             "tuned_us_per_iteration": b.schedule.us_per_iteration, "table_us": b.schedule.table}
    return res


def newton_polish(pot, solver, mu, lam, x, f, mask, steps=3, cg_rtol=1e-12):
    """A few Newton steps on the free DOFs, each solved by the device PCG on the
    exact Hessian-vector pass: FP64 equilibria down to rounding (a PNCG solve
    converges only linearly)."""
    import torch

    for _ in range(steps):
        g = torch.zeros_like(x)
        pot.mesh.eval(pot.material, mu, lam, x, grad=g)
        rhs = (f - g).contiguous()  # -(grad E) = f_ext - internal force
        d = torch.zeros_like(x)
        solver.cg(x, rhs, d, mask, rtol=cg_rtol, max_iters=50000)
        x += d
    return x


def bench_differentiation(dev, n=32, n_fd=64, seed_true=4, seed_init=5, seed_pick=6):
    """Config 4: d/dmu of L = ||x*(mu) - x_target||^2 on an n^3 x 5 cube by the
    adjoint (FP32 forward + adjoint through the autograd path, and FP64),
    checked against FP64 central differences on n_fd random tets.  lambda =
    10 mu (tied, as in the plain large-mesh case); x_target = the FP64
    equilibrium for a seeded ground-truth lognormal mu.  The FP64 equilibria
    of the reference gradient and of the differences are Newton-polished."""
    import torch

    from apple_amd import _lib, adjoint, fem
    from apple_amd.mesh import Region, cube
    from apple_amd.solver import DeviceSolver, PncgConfig, mask_bits

    m = cube(n)
    region = Region.from_mesh(m)
    T = m.n_cells
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8e3 * region.dV / 4, 4))
    fixed = np.zeros((m.n_points, 3), bool)
    fixed[m.points[:, 2] == 0] = True
    mask = mask_bits(fixed, dev)
    mu_true = 1e4 * np.exp(0.5 * np.random.default_rng(seed_true).standard_normal(T))
    mu0 = 1e4 * np.exp(0.5 * np.random.default_rng(seed_init).standard_normal(T))
    m.cell_data["mu"] = mu0
    m.cell_data["lambda"] = 10 * mu0
    pots = {dt: fem.StableNeoHookean.from_region(region, dtype=dt, device=dev) for dt in (torch.float32, torch.float64)}
    p64 = pots[torch.float64]
    s64 = DeviceSolver(p64)
    f64 = torch.as_tensor(f, dtype=torch.float64, device=dev)
    pncg64 = PncgConfig(rtol=1e-10, max_steps=50000, stall_steps=500)

    def equilibrium64(mu_np, x_start):
        mt = torch.as_tensor(mu_np, dtype=torch.float64, device=dev)
        lt = 10 * mt
        x = x_start.clone()
        s64.solve(x, f64, mask, pncg64, mu=mt, lam=lt)
        return newton_polish(p64, s64, mt, lt, x, f64, mask), mt, lt

    t0 = time.perf_counter()
    x_t, _, _ = equilibrium64(mu_true, torch.zeros(m.n_points, 3, dtype=torch.float64, device=dev))
    x_eq, mt0, lt0 = equilibrium64(mu0, torch.zeros_like(x_t))
    # FP64 reference gradient: adjoint on the polished equilibrium
    pa = torch.zeros_like(x_eq)
    s64.cg(x_eq, (-2.0 * (x_eq - x_t)).contiguous(), pa, mask, rtol=1e-12, max_iters=50000, mu=mt0, lam=lt0)
    dmu = torch.empty(T, dtype=torch.float64, device=dev)
    dlam = torch.empty_like(dmu)
    p64.mesh.eval(_lib.APL_SNH, mt0, lt0, x_eq, pa, dmu=dmu, dlam=dlam)
    g_ref = (dmu + 10 * dlam).cpu().numpy()
    log(f"[bench] differentiation: FP64 reference equilibria and adjoint, {time.perf_counter() - t0:.1f} s")

    # the product path: FP32 (plain and refined) and FP64 forward + adjoint through the autograd function
    grads, times = {}, {}
    refine = adjoint.RefineConfig(rtol=1e-10, adjoint_rtol=1e-10, f_ext64=f64)
    variants = (("fp32", torch.float32, PncgConfig(rtol=1e-6, max_steps=50000), adjoint.AdjointConfig(rtol=1e-6), None),
                ("fp32_refined", torch.float32, PncgConfig(rtol=1e-6, max_steps=50000),
                 adjoint.AdjointConfig(rtol=1e-6), refine),
                ("fp64", torch.float64, PncgConfig(rtol=1e-10, max_steps=50000), adjoint.AdjointConfig(rtol=1e-10),
                 None))
    pots[torch.float32].fp64_shadow()  # refinement setup, outside the timings
    for name, dt, cfg, acfg, rf in variants:
        ft = torch.as_tensor(f, dtype=dt, device=dev)
        xt = x_t.to(dt) if rf is None else x_t
        for rep in range(2):  # the first run builds the graphs
            mu_t = torch.as_tensor(mu0, dtype=dt, device=dev).requires_grad_(True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            x = adjoint.equilibrium(pots[dt], mu_t, 10 * mu_t, ft, mask, torch.zeros_like(ft), cfg, acfg, refine=rf)
            ((x - xt) ** 2).sum().backward()
            torch.cuda.synchronize()
            times[name] = time.perf_counter() - t1
        grads[name] = mu_t.grad.double().cpu().numpy()
        if rf is not None:
            fr, ar = adjoint.last_refine_reports()
            refine_steps = {"forward_steps": fr.steps, "forward_cg_iterations": fr.cg_iterations,
                            "adjoint_steps": ar.steps, "adjoint_cg_iterations": ar.cg_iterations}

    # FP64 central differences, warm-started at the polished equilibrium of mu0
    def loss(mu_np):
        xx, _, _ = equilibrium64(mu_np, x_eq)
        return float(((xx - x_t) ** 2).sum())

    pick = np.random.default_rng(seed_pick).choice(T, n_fd, replace=False)
    fd = np.empty(n_fd)
    t0 = time.perf_counter()
    for k, t in enumerate(pick):
        h = 1e-4 * mu0[t]
        mp, mm = mu0.copy(), mu0.copy()
        mp[t] += h
        mm[t] -= h
        fd[k] = (loss(mp) - loss(mm)) / (2 * h)
        if k % 16 == 15:
            log(f"[bench] differentiation: {k + 1} of {n_fd} central differences, {time.perf_counter() - t0:.1f} s")
    t_fd = time.perf_counter() - t0
    out = {"workload": f"{n}^3 x 5 cube ({T} tets), SNH, lambda = 10 mu, mu* lognormal (seed {seed_true}), "
                       f"mu0 lognormal (seed {seed_init}), L = ||x*(mu) - x*(mu*)||^2, {n_fd} FD tets (seed {seed_pick}), "
                       "h = 1e-4 mu, FP64 equilibria Newton-polished",
           "fd_seconds": round(t_fd, 2)}

    def errs(g):
        e = np.abs(g - fd)
        return {"norm_rel_err": float(np.linalg.norm(g - fd) / np.linalg.norm(fd)),
                "max_abs_err_over_max_fd": float(e.max() / np.abs(fd).max())}

    out["fp64_reference_adjoint"] = errs(g_ref[pick])
    for name in ("fp32", "fp32_refined", "fp64"):
        out[name] = {"forward_adjoint_s": round(times[name], 4), **errs(grads[name][pick]),
                     "vs_fp64_reference_norm_rel_err": float(np.linalg.norm(grads[name] - g_ref) / np.linalg.norm(g_ref)),
                     "vs_fp64_reference_max_entry_err_over_max": float(np.abs(grads[name] - g_ref).max()
                                                                       / np.abs(g_ref).max())}
        log(f"[bench] differentiation {name}: ||g - fd|| / ||fd|| = {out[name]['norm_rel_err']:.2e}, "
            f"vs the polished FP64 adjoint {out[name]['vs_fp64_reference_norm_rel_err']:.2e}, {times[name]:.3f} s")
    out["fp32_refined"]["refinement"] = refine_steps
    log(f"[bench] differentiation: polished FP64 adjoint vs FD {out['fp64_reference_adjoint']['norm_rel_err']:.2e}")
    return out


def bench_combined(dev, n_meshes=64, lo=16, hi=48, seed=7, adj_rtol32=1e-6, tight=1e-10):
    """Config 5: forward + adjoint of one batch of uneven meshes with FP32
    element kernels against the same batch run all in FP64 (forward and
    adjoint to rtol 1e-10); per-instance errors of energy, dL/d(mu, lambda)
    and positions, for
      * "fp32": the FP32 path at the reference's tolerance (PNCG and PCG to
        rtol 1e-6, FP64 iterate and reductions);
      * "fp32_refined": the same FP32 solves followed by mixed-precision
        refinement (FP64 residual passes, FP32 PCG corrections) to the FP64
        run's tolerance, forward and adjoint;
      * "fp64_rtol_1e-6": the FP64 run stopped at the FP32 run's tolerance,
        to separate what the tolerance allows from what precision costs.
    L = 1/2 ||x - x0||^2, so the adjoint right-hand side is x0 - x."""
    import torch

    from apple_amd import _lib, batch
    from apple_amd.solver import PncgConfig

    insts, sizes = make_cube_instances(n_meshes, lo, hi, seed, seed)
    res = {}
    diag = np.sqrt(3.0)
    cfg_tight = PncgConfig(rtol=tight, max_steps=50000)
    cfg32 = PncgConfig(rtol=1e-6, max_steps=50000)
    warm = PncgConfig(rtol=1e-6, max_steps=16)

    # all-FP64 reference run
    b64 = batch.Batch(insts, _lib.APL_SNH, dtype=torch.float64, device=dev)
    forward_adjoint(b64, warm, tight, adjoint=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x64, st64, cs64, dmu64, dlam64 = forward_adjoint(b64, cfg_tight, tight)
    torch.cuda.synchronize()
    res["fp64_s"] = round(time.perf_counter() - t0, 3)
    x64, dmu64, dlam64 = x64.double(), dmu64.double(), dlam64.double()
    log(f"[bench] combined FP64: {res['fp64_s']} s")
    x64l, st64l, *_ = forward_adjoint(b64, cfg32, tight, adjoint=False)
    x64l = x64l.double()

    # FP32 at the reference tolerance
    b32 = batch.Batch(insts, _lib.APL_SNH, dtype=torch.float32, device=dev)
    forward_adjoint(b32, warm, adj_rtol32, adjoint=False)
    b32._fp64_copy()  # setup of the refinement's FP64 element data (excluded from the timings)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x32, st32, cs32, dmu32, dlam32 = forward_adjoint(b32, cfg32, adj_rtol32)
    torch.cuda.synchronize()
    res["fp32_s"] = round(time.perf_counter() - t0, 3)
    x32 = x32.double()
    log(f"[bench] combined FP32: {res['fp32_s']} s")

    # FP32 + mixed-precision refinement of the forward and adjoint solutions
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xr, str_ = b32.solve(config=cfg32, refine_fp64="newton", refine_rtol=tight)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    fwd_rep = b32.refine_report
    resolved = list(b32.refine_resolved)
    x64r = b32.x64
    t1 = time.perf_counter()
    pr, csr = b32.adjoint_solve(xr, b32.x0.double() - x64r, rtol=tight, refine_fp64=True,
                                cg_rtol=adj_rtol32)
    dmur = torch.empty(b32.n_tets, dtype=b32.dtype, device=dev)
    dlamr = torch.empty_like(dmur)
    b32.mesh.eval(b32.material, b32.mu, b32.lam, xr, pr, dmu=dmur, dlam=dlamr)
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t0
    adj_rep = b32.adjoint_report
    res["fp32_refined_s"] = round(t_all, 3)
    res["fp32_refined_forward_s"] = round(t_fwd, 3)
    log(f"[bench] combined FP32 + refinement: forward {t_fwd:.3f} s, forward + adjoint {t_all:.3f} s "
        f"({fwd_rep.steps} forward / {adj_rep.steps} adjoint refinement steps, {len(resolved)} FP64 re-solves)")

    def errs(xx, st, dmu, dlam):
        e_err, g_err, g_max, x_err = [], [], [], []
        for i in range(n_meshes):
            sv = slice(int(b32.vtx_off[i]), int(b32.vtx_off[i + 1]))
            stt = slice(int(b32.tet_off[i]), int(b32.tet_off[i + 1]))
            e_err.append(abs(st[i].f - st64[i].f) / max(abs(st64[i].f), 1e-300))
            if dmu is not None:
                gg = torch.cat([dmu[stt].double(), dlam[stt].double()])
                gg64 = torch.cat([dmu64[stt], dlam64[stt]])
                g_err.append(float(torch.linalg.norm(gg - gg64) / torch.linalg.norm(gg64)))
                g_max.append(float((gg - gg64).abs().max() / gg64.abs().max()))
            x_err.append(float((xx[sv] - x64[sv]).abs().max()) / diag)
        return tuple(np.asarray(a) for a in (e_err, g_err, g_max, x_err))

    def summary(e_err, g_err, g_max, x_err):
        out = {"energy_rel_err_max": float(e_err.max()), "pos_err_over_diag_max": float(x_err.max()),
               "pos_err_over_diag_median": float(np.median(x_err)),
               "within_energy_1e-5": int((e_err <= 1e-5).sum()), "within_pos_1e-5_diag": int((x_err <= 1e-5).sum())}
        if len(g_err):
            out.update({"grad_rel_err_max": float(g_err.max()), "grad_rel_err_median": float(np.median(g_err)),
                        "grad_max_entry_err_over_max": float(g_max.max()),
                        "within_grad_1e-4": int((g_err <= 1e-4).sum())})
        return out

    E32 = errs(x32, st32, dmu32, dlam32)
    ER = errs(x64r, str_, dmur, dlamr)
    EL = errs(x64l, st64l, None, None)
    # branch switches: the refined FP32 answer is an FP64 equilibrium (residual at the
    # target) but not the FP64 run's one
    switched = [i for i in range(n_meshes) if ER[3][i] > 1e-5 and fwd_rep.converged[i] and i not in resolved]
    on_branch = np.array([i not in switched for i in range(n_meshes)])
    res.update({
        "workload": f"{n_meshes} cubes, n ~ U{{{lo}..{hi}}} (seed {seed}), {b32.n_tets} tets, SNH, per-instance "
                    f"mu/nu as the batched case; L = 1/2 ||x - x0||^2; FP64 run to rtol {tight:g}",
        "fp32_converged": int(sum(s.success for s in st32)),
        "fp32_stagnated": int(sum(s.result.name == "STAGNATION" for s in st32)),
        "fp64_converged": int(sum(s.success for s in st64)),
        "fp32": summary(*E32),
        "fp32_refined": {
            **summary(*ER),
            "forward_refinement_steps": fwd_rep.steps,
            "forward_refinement_cg_iterations": fwd_rep.cg_iterations,
            "adjoint_refinement_steps": adj_rep.steps,
            "adjoint_refinement_cg_iterations": adj_rep.cg_iterations,
            "fp64_pncg_resolved": resolved,
            "forward_residual_converged": int(fwd_rep.converged.sum()),
            "adjoint_residual_converged": int(adj_rep.converged.sum()),
            "on_fp64_branch": int(on_branch.sum()),
            "on_branch_energy_max": float(ER[0][on_branch].max()),
            "on_branch_pos_max": float(ER[3][on_branch].max()),
            "on_branch_grad_max": float(ER[1][on_branch].max()),
            "branch_switches": [{"instance": i, "n": sizes[i], "pos_err_over_diag": float(ER[3][i]),
                                 "energy_gap_rel": (str_[i].f - st64[i].f) / abs(st64[i].f),
                                 "fp32_result": st32[i].result.name} for i in switched],
        },
        "fp64_rtol_1e-6": summary(*EL),
        "per_instance": [{"n": sizes[i], "fp32": st32[i].result.name, "e": float(E32[0][i]), "g": float(E32[1][i]),
                          "x": float(E32[3][i]), "x_refined": float(ER[3][i]), "e_refined": float(ER[0][i]),
                          "g_refined": float(ER[1][i]), "x_fp64_rtol_1e-6": float(EL[3][i])}
                         for i in range(n_meshes)],
    })
    r = res["fp32_refined"]
    log(f"[bench] combined: FP32 at rtol 1e-6: positions within 1e-5 {res['fp32']['within_pos_1e-5_diag']}/{n_meshes}, "
        f"dL/dq within 1e-4 {res['fp32']['within_grad_1e-4']}/{n_meshes}; refined: energy max {r['energy_rel_err_max']:.2e}, "
        f"positions within 1e-5 {r['within_pos_1e-5_diag']}/{n_meshes}, dL/dq within 1e-4 {r['within_grad_1e-4']}/{n_meshes}; "
        f"on the FP64 branch ({r['on_fp64_branch']}): energy {r['on_branch_energy_max']:.2e}, positions "
        f"{r['on_branch_pos_max']:.2e}, dL/dq {r['on_branch_grad_max']:.2e}; branch switches {switched}")
    return res


def bench_config1(dev, cpu_threads):
    """The CPU configuration solved end to end: 10^3 x 5 cube, SNH mu = 1e4,
    lambda = 1e5, rho = 1e3, lumped gravity, z = 0 face fixed, u0 ~ N(0, 0.01)
    on the free DOFs (seed 0), PNCG to 1e-6.  CPU: the C restatement's PNCG
    (FP64, separate passes with scatter) on 1 and cpu_threads threads; GPU:
    the device PNCG in FP32 and FP64 (mesh and graph set up beforehand)."""
    import torch

    from apple_amd import fem
    from apple_amd.mesh import Region, cube
    from apple_amd.solver import DeviceSolver, PncgConfig, mask_bits
    from oracle import c_oracle

    m = cube(10)
    m.cell_data["mu"] = np.full(m.n_cells, 1e4)
    m.cell_data["lambda"] = np.full(m.n_cells, 1e5)
    region = Region.from_mesh(m)
    f = np.zeros((m.n_points, 3))
    np.add.at(f[:, 2], m.cells.reshape(-1), np.repeat(-9.8 * 1e3 * region.dV / 4.0, 4))
    fixed = np.zeros((m.n_points, 3), bool)
    fixed[m.points[:, 2] == 0.0] = True
    x0 = 0.01 * np.random.default_rng(0).standard_normal((m.n_points, 3))
    x0[fixed] = 0.0
    out = {"workload": "10^3 x 5 cube (5000 tets), SNH mu 1e4 lambda 1e5, gravity rho 1e3, z = 0 fixed, "
                       "u0 ~ N(0, 0.01) seed 0, PNCG to 1e-6 relative gradient"}
    cp = c_oracle.SnhProblem(m.cells, region.dhdX, region.dV, m.cell_data["mu"], m.cell_data["lambda"])
    for threads in (1, cpu_threads):
        used = c_oracle.set_threads(threads)
        t0 = time.perf_counter()
        _, steps, status, _ = cp.pncg(x0, f, ~fixed.reshape(-1), rtol=1e-6, max_steps=20000)
        dt = time.perf_counter() - t0
        out[f"cpu_{used}_threads"] = {"s": round(dt, 4), "steps": int(steps), "status": int(status)}
    for dtype in (torch.float32, torch.float64):
        pot = fem.StableNeoHookean.from_region(region, dtype=dtype, device=dev)
        solver = DeviceSolver(pot)
        t = lambda a: torch.as_tensor(a, dtype=dtype, device=dev).contiguous()
        mask = mask_bits(fixed, dev)
        xw = t(x0)
        solver.solve(xw, t(f), mask, PncgConfig(rtol=1e-6, max_steps=16))  # graph capture
        x = t(x0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve(x, t(f), mask, PncgConfig(rtol=1e-6, max_steps=20000))[0]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[f"gpu_{str(dtype).split('.')[-1]}"] = {"s": round(dt, 4), "steps": st.n_steps, "status": st.result.name}
    log(f"[bench] config 1 end to end: CPU 1 thread {out['cpu_1_threads']['s']} s, GPU FP32 "
        f"{out['gpu_float32']['s']} s, FP64 {out['gpu_float64']['s']} s")
    return out


def bench_setup(dev, n):
    """Mesh setup of the n^3 x 5 cube: the host region (numpy inverse / det,
    as the reference's Region) + apl_mesh_create, against
    apl_mesh_create_from_points (element data by a device kernel), with the
    block partition on the device (default) and on the host; FP64 meshes
    (whose slot cap makes the greedy block ends bind); the Morton order of
    the cube with its tets and vertices shuffled, device vs numpy; lumped
    vertex volumes, device vs numpy's np.add.at."""
    import torch

    from apple_amd import _lib
    from apple_amd.mesh import Mesh, Region, cube, reorder, reorder_host
    from apple_amd.runtime import DeviceMesh, layout_mode

    m = cube(n)
    DeviceMesh(m.cells, None, None, m.n_points, dtype=torch.float32, device=dev, points=m.points).close()  # warm

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    r, t_region = timed(lambda: Region.from_mesh(m))
    a, t_host_create = timed(lambda: DeviceMesh(r.cells, r.dminv, r.dV, m.n_points, dtype=torch.float32, device=dev))
    a.close()
    out = {"n_tets": m.n_cells, "host_region_s": round(t_region, 4), "host_mesh_create_s": round(t_host_create, 4)}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        b, t_dev = timed(lambda: DeviceMesh(m.cells, None, None, m.n_points, dtype=dt, device=dev, points=m.points))
        b.close()
        prev = layout_mode(_lib.LAYOUT_HOST)
        try:
            c, t_hp = timed(lambda: DeviceMesh(m.cells, None, None, m.n_points, dtype=dt, device=dev,
                                               points=m.points))
        finally:
            layout_mode(prev)
        c.close()
        out[f"device_setup_{name}_s"] = round(t_dev, 4)
        out[f"device_setup_{name}_host_partition_s"] = round(t_hp, 4)
    out["device_setup_s"] = out["device_setup_fp32_s"]
    rng = np.random.default_rng(0)
    vp = rng.permutation(m.n_points)
    rank = np.empty_like(vp)
    rank[vp] = np.arange(vp.size)
    shuffled = Mesh(m.points[vp], rank[m.cells][rng.permutation(m.n_cells)].astype(np.int32))
    reorder(shuffled, dev)  # warm
    (_, vpd, _), t_rd = timed(lambda: reorder(shuffled, dev))
    (_, vph, _), t_rh = timed(lambda: reorder_host(shuffled))
    out["morton_reorder_device_s"] = round(t_rd, 4)
    out["morton_reorder_numpy_s"] = round(t_rh, 4)
    out["morton_reorder_equal"] = bool(np.array_equal(vpd, vph))
    b = DeviceMesh(m.cells, None, None, m.n_points, dtype=torch.float64, device=dev, points=m.points)
    b.lumped_volume()
    vol, t_ld = timed(lambda: b.lumped_volume())
    dV = b.region()[1].cpu().numpy()

    def host_lumped():
        acc = np.zeros(m.n_points)
        np.add.at(acc, m.cells.reshape(-1), np.repeat(dV / 4.0, 4))
        return acc

    ref, t_lh = timed(host_lumped)
    b.close()
    out["lumped_volume_device_s"] = round(t_ld, 5)
    out["lumped_volume_numpy_s"] = round(t_lh, 4)
    out["lumped_volume_rel_err"] = float(np.abs(vol.cpu().numpy() - ref).max() / np.abs(ref).max())
    log(f"[bench] setup {n}^3: host region {t_region:.3f} s + mesh {t_host_create:.3f} s; device setup FP32 "
        f"{out['device_setup_fp32_s']:.3f} s (host partition {out['device_setup_fp32_host_partition_s']:.3f} s), "
        f"FP64 {out['device_setup_fp64_s']:.3f} s (host partition {out['device_setup_fp64_host_partition_s']:.3f} s); "
        f"Morton reorder device {t_rd:.3f} s vs numpy {t_rh:.3f} s; lumped volumes {t_ld * 1e3:.2f} ms vs "
        f"{t_lh * 1e3:.1f} ms")
    return out


def bench_inverse(dev, n=32, steps=12, prestrain=0.9, lr=0.05, seed=9):
    """Muscle-activation inverse loop (FP32): per-tet log activation, surface
    point-to-point loss against a forward solution for a smooth true field,
    Adam; every step = warm-started PNCG + adjoint PCG + mixed pass."""
    import torch

    from apple_amd import fem, inverse
    from apple_amd.mesh import Region, cube
    from apple_amd.solver import DeviceSolver, PncgConfig, mask_bits

    m = cube(n)
    rng = np.random.default_rng(seed)
    mu = 1e4 * np.exp(0.3 * rng.standard_normal(m.n_cells))
    m.cell_data["mu"] = mu
    m.cell_data["lambda"] = 10 * mu
    m.cell_data["activation"] = np.zeros((m.n_cells, 6))
    region = Region.from_mesh(m)
    pot = fem.StableNeoHookeanMuscle.from_region(region, dtype=torch.float32, device=dev)
    fixed = np.zeros((m.n_points, 3), bool)
    fixed[m.points[:, 0] == 0] = True
    mask = mask_bits(fixed, dev)
    surface = np.flatnonzero((m.points == 0).any(1) | (m.points == 1).any(1))
    centroid = m.points[m.cells].mean(1)
    T = m.n_cells
    true = torch.as_tensor(0.2 + 0.1 * np.sin(2 * np.pi * centroid[:, 1]), dtype=torch.float32, device=dev)
    mk = lambda p: {"activation": inverse.muscle_x_activation(p, prestrain=prestrain).contiguous()}
    x_t = torch.zeros(m.n_points, 3, dtype=torch.float32, device=dev)
    DeviceSolver(pot).solve(x_t, None, mask, PncgConfig(rtol=1e-6, max_steps=50000), activation=mk(true)["activation"])
    idx = torch.as_tensor(surface, device=dev)
    inv = inverse.Inverse(pot, [inverse.PointToPointLoss(idx, x_t[idx])], mk, mask=mask, lr=lr,
                          forward=PncgConfig(rtol=1e-6, max_steps=50000))
    log(f"[bench] inverse: {T} tets, {len(surface)} surface targets, {steps} Adam steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p = inv.solve(torch.zeros(T, dtype=torch.float32, device=dev), steps=steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    h = inv.history
    err0 = float(true.abs().mean())
    err = float((p - true).abs().mean())
    log(f"[bench] inverse: {dt / steps * 1e3:.1f} ms per step, loss {h[0].loss:.3e} -> {h[-1].loss:.3e}, "
        f"mean |p - p_true| {err0:.3f} -> {err:.3f}")
    return {
        "workload": f"muscle activation inverse, {n}^3 cube x 5 tets ({T} per-tet parameters), FP32, "
        f"surface point-to-point loss, Adam lr {lr}",
        "steps": steps,
        "ms_per_step": round(dt / steps * 1e3, 2),
        "first_step_ms": round(h[0].seconds * 1e3, 2),
        "forward_steps_first": h[0].forward_steps,
        "forward_steps_mean_warm": round(float(np.mean([r.forward_steps for r in h[1:]])), 1),
        "adjoint_iters_mean": round(float(np.mean([r.adjoint_iters for r in h])), 1),
        "loss_first": h[0].loss,
        "loss_last": h[-1].loss,
        "param_mean_abs_err_start": round(err0, 4),
        "param_mean_abs_err_end": round(err, 4),
    }


if __name__ == "__main__":
    main()
