#!/usr/bin/env python3
"""Benchmark: compressible viscous RHS on one 256^3 FP64 subdomain (MI355X).

Workload: 256^3 FP64, 5 conserved variables, uniform flow at Mach 0.5 in x
(gamma 1.4, mu = 1/Re = 1e-4), seeded Gaussian acoustic pulse (amplitude
1e-3, width 8 cells), non-reflective characteristic BCs on all six faces.
One timed step = one RK4 step: time step, four stages (each = full RHS
through the HIP library + stage update) and the compact filter of the five
conserved variables along x, y, z.  value = RHS grid points per second =
4 * 256^3 / t_step (filter and time step included, i.e. conservative).

Prints ONE JSON line.  --gpus N > 1 is not supported: this build runs the
subdomain of one GPU.  The default run also measures the other workloads
(--workload plain64, batch96, precision, adjoint, curv256), each in a child
process of this script, and carries their lines under "other_workloads"
(--no-workloads skips them).
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="grid points per axis")
    ap.add_argument("--path", default="fused", choices=["fused", "ref"],
                    help="fused sweeps (default) or the reference-ordered per-operator path")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only launch the dominant sweep and a plain device copy --reps times each "
                         "(the program rocprofv3 --pmc runs for the traffic figures)")
    ap.add_argument("--adjoint-probe", action="store_true",
                    help="only run --reps RK4 steps recording tapes and their adjoints on the 128^3 Euler "
                         "workload of --workload adjoint (the program rocprofv3 profiles the adjoint kernels with)")
    ap.add_argument("--batch-probe", action="store_true",
                    help="only run --reps steps of the 96-subdomain batch in the tuned grid mode (one launch per "
                         "subdomain and operator) and a plain device copy of known size (the program rocprofv3 "
                         "profiles the batched kernels with)")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--workload", default="rhs", choices=["rhs", "plain64", "batch96", "precision", "adjoint", "curv256"],
                    help="rhs: the 256^3 viscous RK4 step (default); plain64: derivative x/y/z + filter x/y/z "
                         "of a periodic 64^3 block, against the analytic derivative and the CPU oracle; batch96: "
                         "derivative x/y/z + filter "
                         "x/y/z of 96 periodic subdomains with seeded sizes in [48, 320]; precision: 50 RK4 "
                         "steps at 512x512x256 with FP64 and with FP32+refinement line solves; adjoint: gradient "
                         "of an acoustic-energy loss after 40 Euler RK4 steps at 128^3 (checkpointed adjoint); curv256: the "
                         "RHS workload on a skewed curvilinear 256^3 grid, fused sweeps with per-point metrics "
                         "against the per-operator path")
    ap.add_argument("--curvilinear", action="store_true",
                    help="--workload adjoint on the sine-skewed curvilinear grid (per-point metrics)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-workloads", action="store_true",
                    help="default workload only: do not run the other workloads' lines")
    ap.add_argument("--cpu-reps", type=int, default=3)
    return ap.parse_args()


def built():
    return os.path.exists(os.path.join(ROOT, "compact_schemes_amd", "libcompact_schemes_amd.so"))


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs one GPU's subdomain only; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if not built():
        import __graft_entry__
        __graft_entry__.build()
    import numpy as np
    import torch
    import compact_schemes_amd as cs
    from compact_schemes_amd import workloads as wl

    torch.cuda.set_device(0)
    if args.pmc_probe:
        return pmc_probe(args)
    if args.adjoint_probe:
        return adjoint_probe(args)
    if args.batch_probe:
        return batch_probe(args)
    if args.workload == "plain64":
        return plain_bench(args)
    if args.workload == "batch96":
        return batch_bench(args)
    if args.workload == "precision":
        return precision_bench(args)
    if args.workload == "adjoint":
        return adjoint_bench(args)
    if args.workload == "curv256":
        return curvilinear_bench(args)
    n = args.n
    dims = (n, n, n)
    lmx = n ** 3
    st = wl.flow_state(dims, seed=2)
    q = torch.from_numpy(st["qa"]).cuda()
    q_init = q.clone()

    if args.path == "ref":
        from compact_schemes_amd.stepper import Stepper
        stepper = Stepper(dims, wl.NBC_NON_REFLECTIVE, True, st["umf"], st)
        step = lambda: stepper.rk4_step(q)
        rhs = lambda: stepper.rhs(q)
        num = stepper.num
    else:
        from compact_schemes_amd.fused import FILTER_SWEEPS, FusedSolver
        solver = FusedSolver(dims, wl.NBC_NON_REFLECTIVE, st)
        step = lambda: solver.rk4_step(q)
        rhs = lambda: solver.rhs(q)
        num = solver.num

    def run_timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        q.copy_(q_init)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0
    t_eager = run_timed(step)
    t_total = t_eager
    timing = "eager launches"
    if args.path == "fused":
        # the whole RK4 step (about 30 launches, no host synchronisation) as one
        # hipGraph replayed per step
        gs = torch.cuda.Stream()
        graph = torch.cuda.CUDAGraph()
        q.copy_(q_init)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=gs):
            solver.rk4_step(q, stream=gs)
        t_total = run_timed(graph.replay)
        timing = "hipGraph replay of one RK4 step"
    ms_step = 1e3 * t_total / args.steps
    value = 4.0 * lmx * args.steps / t_total / 1e9
    assert torch.isfinite(q).all().item(), "non-finite state after the timed steps"

    # RHS alone (HIP events on the launch stream)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    q.copy_(q_init)
    rhs()
    e0.record(stream)
    nr = 5
    for _ in range(nr):
        rhs()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_rhs = e0.elapsed_time(e1) / nr

    # dominant kernel: the y-direction line sweep, timed in isolation
    kern = {}
    if args.path == "ref":
        fin = q[0:lmx].clone()
        fout = torch.empty_like(fin)
        e0.record(stream)
        nk = 20
        for _ in range(nk):
            num.deriv(1, fin, fout)
        e1.record(stream)
        torch.cuda.synchronize()
        ms_k = e0.elapsed_time(e1) / nk
        bytes_per_launch = 16.0 * lmx          # read + write one FP64 field
        kern = {"name": "cs::line_kernel<16,0,1> (y derivative)", "ms": ms_k, "bytes": bytes_per_launch}
    else:
        q.copy_(q_init)
        kern = solver.time_dominant_kernel(q, stream)
        kern["traffic"] = traffic_from_profile(kern["kernel"], n)
        if solver.filter_fused:                   # the three filter sweeps that close each step
            kern["all"] += [{"sweep": r["name"], "kernel": r["kernel"], "ms": round(r["ms"], 5),
                             "GBs": round(r["bytes"] / (r["ms"] * 1e-3) / 1e9, 1)}
                            for r in solver.time_sweeps(q, stream, reps=5, sweeps=FILTER_SWEEPS)]

    achieved = kern["bytes"] / (kern["ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": kern.get("traffic"),
                "kernel": kern.get("kernel", kern["name"]), "sweep": kern["name"],
                "kernel_ms": round(kern["ms"], 5), "algorithmic_bytes_per_launch": kern["bytes"]}
    if "all" in kern:
        roofline["sweeps"] = kern["all"]
    fl = fp64_flops_from_profile(roofline["kernel"], n) if args.path == "fused" else None
    if roofline.get("traffic") is not None or fl is not None:
        # the counter figures come from the committed summaries: say whether
        # they were measured with these kernel sources
        roofline["profile_matches_sources"] = (profile_digest(PROFILE_TRAFFIC) == source_digest() and
                                               profile_digest(PROFILE_FP64) == source_digest())
    if fl is not None:
        # the second roof: FP64 vector rate of the same kernel (counter flops / HIP-event time)
        ach = fl / (kern["ms"] * 1e-3) / 1e12
        roofline["fp64"] = {"achieved": round(ach, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": round(ach / FP64_PEAK_TFLOPS, 4), "flops_per_launch": fl}

    if args.path == "fused":
        # the whole path against the same roof: credited bytes of one RHS
        # (736 B/pt) and of one RK4 step (4 stages of 776 B/pt, three filter
        # sweeps of 80) over their measured times
        rhs_gbs = RHS_BYTES_PER_PT * lmx / (ms_rhs * 1e-3) / 1e9
        step_gbs = STEP_BYTES_PER_PT * lmx / (ms_step * 1e-3) / 1e9
        roofline["whole_rhs"] = {"bytes_per_pt": RHS_BYTES_PER_PT, "ms": round(ms_rhs, 4),
                                 "achieved": round(rhs_gbs, 1), "frac": round(rhs_gbs / HBM_PEAK_GBS, 4)}
        roofline["whole_rk4_step"] = {"bytes_per_pt": STEP_BYTES_PER_PT, "ms": round(ms_step, 4),
                                      "achieved": round(step_gbs, 1), "frac": round(step_gbs / HBM_PEAK_GBS, 4)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(st, dims, args.cpu_reps)

    line = {
        "metric": "RHS Gpts/s at 256^3 FP64",
        "value": round(value, 4),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded acoustic pulse on uniform Mach 0.5 flow)",
        "config": {"workload": f"viscous compressible RHS, {n}^3 FP64, non-reflective gcbc on all faces, "
                               "one RK4 step (4 RHS + stage updates + GV10 filter)",
                   "grid": [n, n, n], "path": args.path, "timing": timing,
                   "ms_per_step_eager": round(1e3 * t_eager / args.steps, 3), "ms_per_rhs": round(ms_rhs, 3),
                   "rhs_gpts_s_rhs_only": round(lmx / (ms_rhs * 1e-3) / 1e9, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if not args.no_workloads and args.path == "fused" and n == 256:
        line["other_workloads"] = other_workloads()
    print(json.dumps(line), flush=True)


OTHER_WORKLOADS = ("plain64", "batch96", "precision", "adjoint", "curv256")


def other_workloads(timeout=300):
    """The JSON lines of the other workloads, each from a child process of
    this script (its own GPU context and memory; a failure or a time-out is
    recorded, not raised)."""
    out = {}
    for w in OTHER_WORKLOADS:
        t0 = time.perf_counter()
        try:
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--workload", w, "--no-cpu"],
                               capture_output=True, text=True, timeout=timeout, cwd=ROOT)
            lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
            if r.returncode == 0 and lines:
                out[w] = json.loads(lines[-1])
            else:
                out[w] = {"error": f"exit status {r.returncode}", "stderr": r.stderr[-400:]}
        except subprocess.TimeoutExpired:
            out[w] = {"error": f"no line within {timeout} s"}
        print(f"workload {w}: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return out


PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic_rk_stage_256.txt")
RHS_BYTES_PER_PT = 736                     # six sweeps of one viscous RHS
# RK4 step: four stages and the filter along x, y, z; the time step's maxima
# are reduced inside the first gradient sweep of the first stage (the y sweep
# of the 256^3 workload: 16 lines x 32 chunks, whole waves), which reads q anyway
STEP_BYTES_PER_PT = 4 * 776 + 3 * 80
PROFILE_FP64 = os.path.join(ROOT, "profiles", "pmc_fp64_rk_stage_256.txt")


# credited bytes of one curvilinear RHS: the 736 of the constant-metric RHS plus
# the closure's nine metrics and Jacobian (80), each flux sweep's metric row
# (24) and the five stress and heat-flux components beyond its own row (40)
CURV_RHS_BYTES_PER_PT = RHS_BYTES_PER_PT + 80 + 3 * (24 + 40)


def curvilinear_bench(args, n=256):
    """The RHS workload (viscous, Mach 0.5 with the acoustic pulse,
    non-reflective faces) on the skewed grid x = i + 1.5 sin(2 pi j/n), ...
    with its analytic metrics: one RK4 step and one RHS of the fused solver
    with per-point metrics (cs_solver_set_metrics), and one RHS of the
    per-operator path (calc_viscous_shear_stress + calc_fluxes + gcbc_go of
    the same arrays).  value = RHS grid points per second per RK4 step."""
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    from compact_schemes_amd.stepper import Stepper
    dims = (n, n, n)
    lmx = n ** 3
    st = wl.curvilinear_state(dims, seed=2, skew=1.5)
    q0 = torch.from_numpy(st["qa"]).cuda()
    q = q0.clone()
    sv = FusedSolver(dims, wl.NBC_NON_REFLECTIVE, st, curvilinear=True)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        q.copy_(q0)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps
    ms_step = timed(lambda: sv.rk4_step(q), args.steps)
    assert torch.isfinite(q).all().item(), "non-finite state after the timed steps"
    ms_rhs = timed(lambda: sv.rhs(q), 5)
    ref = Stepper(dims, wl.NBC_NON_REFLECTIVE, True, st["umf"], st)
    ms_ref = timed(lambda: ref.rhs(q), 3)
    gbs = CURV_RHS_BYTES_PER_PT * lmx / (ms_rhs * 1e-3) / 1e9
    line = {
        "metric": "curvilinear RHS Gpts/s at 256^3 FP64",
        "value": round(4.0 * lmx / (ms_step * 1e-3) / 1e9, 4),
        "unit": "Gpts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded acoustic pulse on uniform Mach 0.5 flow, sine-skewed grid, analytic metrics)",
        "config": {"workload": "viscous compressible RHS, 256^3 FP64 curvilinear grid, non-reflective gcbc on all "
                               "faces, one RK4 step (4 RHS + stage updates + GV10 filter)",
                   "grid": [n, n, n], "path": "fused, per-point metrics", "timing": "eager launches",
                   "ms_per_rhs": round(ms_rhs, 3), "rhs_gpts_s_rhs_only": round(lmx / (ms_rhs * 1e-3) / 1e9, 4),
                   "per_operator_ms_per_rhs": round(ms_ref, 3),
                   "fused_speedup_over_per_operator": round(ms_ref / ms_rhs, 3)},
        "roofline": {"whole_rhs": {"bytes_per_pt": CURV_RHS_BYTES_PER_PT, "ms": round(ms_rhs, 4),
                                   "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}},
    }
    print(json.dumps(line), flush=True)


def plain_bench(args, n=64):
    """Config 1: one periodic 64^3 block, f = 8 seeded plane waves (wavenumbers
    1..20) + 1e-3 noise; one step = d/dx, d/dy, d/dz of f, then the GV10
    filter (alpha_f = 0.49) of f along x, y, z.  Results against the oracle
    (the reference's algorithm on the CPU) and the derivatives of the
    noise-free field against the analytic ones; the oracle's own time for
    the same step on the host cores beside it."""
    import numpy as np
    import torch
    import compact_schemes_amd as cs
    from compact_schemes_amd import workloads as wl
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    dims = (n, n, n)
    per = wl.NBC_PERIODIC
    f, exact = wl.periodic_field(dims, seed=1, modes=8, kmax=20, noise=1e-3)
    fclean, _ = wl.periodic_field(dims, seed=1, modes=8, kmax=20, noise=0.0)
    num = cs.Numerics(dims, per)
    fd = torch.from_numpy(f).cuda()
    d = [torch.empty_like(fd) for _ in range(3)]
    g = torch.empty_like(fd)

    st = torch.cuda.Stream()

    def step():
        for a in range(3):
            num.deriv(a, fd, d[a], stream=st)
        num.filter(0, fd, g, stream=st)
        num.filter(1, g, g, stream=st)
        num.filter(2, g, g, stream=st)

    # a step is tens of microseconds: time many of them with HIP events on the
    # launch stream
    reps = max(args.steps, 500)

    def timed(fn):
        with torch.cuda.stream(st):
            for _ in range(args.warmup):
                fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                fn()
            e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps
    t_eager = timed(step)
    # the six launches are a few microseconds of work each at 64^3: replay
    # them from a hipGraph (no allocation or memset inside the step); a replay
    # runs on torch's current stream, st inside timed()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=st):
        step()
    t_graph = timed(graph.replay)
    # the value is the faster of the two ways to run the step (a replay adds a
    # few microseconds at this size)
    t_gpu = min(t_eager, t_graph)
    timing = "eager launches on one stream" if t_eager <= t_graph else "hipGraph replay of the six launches"
    ctx = pyoracle.Context(dims, per)
    cores = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "1024")))
    pyoracle.set_num_threads(cores)
    t0 = time.perf_counter()
    ref_d = [ctx.deriv_field(a, f) for a in range(3)]
    ref_g = ctx.filter(2, ctx.filter(1, ctx.filter(0, f)))
    t_cpu = time.perf_counter() - t0
    nrel = lambda a, b: float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    parity = max([nrel(d[a].cpu().numpy(), ref_d[a]) for a in range(3)] + [nrel(g.cpu().numpy(), ref_g)])
    dclean = torch.empty_like(fd)
    analytic = []
    for a in range(3):
        num.deriv(a, torch.from_numpy(fclean).cuda(), dclean)
        analytic.append(float(np.max(np.abs(dclean.cpu().numpy() - exact[a]))))
    npts = n ** 3
    line = {
        "metric": "plain method Gpts/s (periodic 64^3 FP64 block)",
        "value": round(npts / t_gpu / 1e9, 4), "unit": "Gpts/s", "n_gpus": 1, "steps": reps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_gpu, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (8 seeded plane waves, wavenumbers 1..20, + 1e-3 noise)",
        "config": {"workload": "d/dx, d/dy, d/dz + GV10 filter x, y, z, periodic 64^3",
                   "timing": timing, "ms_per_step_eager": round(1e3 * t_eager, 4),
                   "ms_per_step_graph": round(1e3 * t_graph, 4),
                   "parity_vs_oracle": parity,
                   "abs_err_vs_analytic_xyz": analytic,
                   "note": "the analytic error is the scheme's own (modified wavenumber at k up to 20 of 64, "
                           "and the as-shipped seam rows)"},
        "cpu_baseline": {"value": round(npts / t_cpu / 1e9, 5), "unit": "Gpts/s", "cores": pyoracle.num_threads(),
                         "kind": "port", "sample": "one step of the oracle (3 derivatives + 3 filter passes) at 64^3"},
    }
    print(json.dumps(line), flush=True)


def batch_bench(args):
    """Config 3: the plain method on 96 uneven periodic subdomains: one step =
    derivative along x, y, z + filter along x, y, z, replayed from a hipGraph
    of the batched launches, after the autotuner has picked chunk lengths and
    the grid mode (one workgroup per item, persistent, or one launch per
    subdomain and operator).  The untuned batch in each grid mode and the same
    work done subdomain by subdomain through cs_deriv / cs_filter (not
    captured) are timed beside it.  value = grid points of all subdomains per second per step."""
    import torch
    import compact_schemes_amd as cs
    from compact_schemes_amd import batch as bm
    per = [[45, 45, 45], [45, 45, 45]]
    from compact_schemes_amd.workloads import batch_fields
    dims, f0 = batch_fields()
    npts = sum(d[0] * d[1] * d[2] for d in dims)
    f = [t.clone() for t in f0]
    d = [torch.empty_like(f[s // 3]) for s in range(3 * len(dims))]
    st = torch.cuda.Stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, n):
        for _ in range(args.warmup):
            fn()
        st.synchronize()
        e0.record(st)
        for _ in range(n):
            fn()
        e1.record(st)
        st.synchronize()
        return e0.elapsed_time(e1) / n

    bm.tune_clear()
    b = bm.Batch(dims, per)
    b.bind(f, d)
    ms_default = timed(lambda: b.step(st), args.steps)
    b.set_persistent(1)
    ms_persistent = timed(lambda: b.step(st), args.steps)
    b.set_persistent(2)
    ms_per_subdomain = timed(lambda: b.step(st), args.steps)
    b.set_persistent(0)
    nums = [cs.Numerics(dm, per) for dm in dims]

    def one_at_a_time():
        for s, num in enumerate(nums):
            for axis in range(3):
                num.deriv(axis, f[s], d[3 * s + axis], stream=st)
            for axis in range(3):
                num.filter(axis, f[s], f[s], stream=st)
    ms_single = timed(one_at_a_time, args.steps)
    b.autotune(reps=3, stream=st)
    b.capture(st)
    torch.cuda.synchronize()
    ms_graph = timed(lambda: b.replay(st), args.steps)
    items, launches = b.info()
    mode = b.get_persistent()
    bm.tune_clear()
    line = {
        "metric": "batched plain method Gpts/s (96 uneven FP64 subdomains)",
        "value": round(npts / (ms_graph * 1e-3) / 1e9, 4),
        "unit": "Gpts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_graph, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (8 seeded plane waves + 1e-3 noise per subdomain, seeds 3..98)",
        "config": {"workload": "derivative x, y, z + GV10 filter x, y, z on 96 periodic subdomains, sizes in [48, 320]",
                   "points": npts, "tuned_graph_ms": round(ms_graph, 3),
                   "untuned_batched_ms": round(ms_default, 3), "untuned_persistent_ms": round(ms_persistent, 3),
                   "untuned_per_subdomain_launches_ms": round(ms_per_subdomain, 3),
                   "one_subdomain_at_a_time_ms": round(ms_single, 3),
                   "tuned_grid_mode": {0: "one workgroup per item", 1: "persistent",
                                       2: "one launch per subdomain and operator"}[mode],
                   "items": items, "launches": launches,
                   "algorithmic_GBs": round(npts * 96 / (ms_graph * 1e-3) / 1e9, 1)},
    }
    print(json.dumps(line), flush=True)


def precision_bench(args, dims=(512, 512, 256), nsteps=50):
    """Config 4: the flow of the RHS workload on 512x512x256 with periodic
    faces, 50 RK4 steps (viscous, filter every step, dt fixed from the CFL
    limit of the initial state).  Primary: the fused solver with FP64
    intermediates against the same solver with the viscous intermediates
    (gradients, stress, heat flux) stored in FP32 (cs_solver_set_precision
    1: 588 instead of 736 HBM bytes per point and RHS).  Secondary: the
    reference-ordered operators with FP64 line solves against FP32 solves +
    one FP64-residual refinement sweep (cs_set_precision 1).  Errors per
    conserved variable relative to the state and to its change over the run;
    value = the largest state-relative error of the fused FP32 run."""
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    from compact_schemes_amd.stepper import Stepper
    per = wl.NBC_PERIODIC
    st = wl.flow_state(dims, seed=4)
    lmx = dims[0] * dims[1] * dims[2]
    q0 = torch.from_numpy(st["qa"]).cuda()
    del st["qa"]

    def run(make, tag):
        stepper = make()
        q = q0.clone()
        stepper.rk4_step(q, dt=dt)                 # warm-up step, then restart
        q.copy_(q0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(nsteps):
            stepper.rk4_step(q, dt=dt)
            if (k + 1) % 10 == 0:
                torch.cuda.synchronize()
                print(f"[precision] {tag}: {k + 1} steps", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / nsteps
        del stepper
        torch.cuda.empty_cache()
        return q, t

    def fused(prec):
        sv = FusedSolver(dims, per, st)
        sv.set_precision(prec)
        return sv

    def ref_ordered(mode):
        stp = Stepper(dims, per, True, st["umf"], st)
        stp.num.set_precision(mode)
        return stp

    dt = FusedSolver(dims, per, st).time_step(q0)
    f64, t_f64 = run(lambda: fused(0), "fused fp64")
    f32, t_f32 = run(lambda: fused(1), "fused fp32 intermediates")
    r64, t_r64 = run(lambda: ref_ordered(0), "per-operator fp64")
    r32, t_r32 = run(lambda: ref_ordered(1), "per-operator fp32 solves + refinement")

    def errors(a_all, b_all):
        err = []
        for v in range(5):
            sl = slice(v * lmx, (v + 1) * lmx)
            a, b, c0 = a_all[sl], b_all[sl], q0[sl]
            d = a - b
            ch = b - c0
            err.append({"var": ["rho", "rho_u", "rho_v", "rho_w", "E"][v],
                        "max_rel_state": float(d.abs().max() / b.abs().max()),
                        "l2_rel_state": float(d.norm() / b.norm()),
                        "max_rel_change": float(d.abs().max() / ch.abs().max()),
                        "l2_rel_change": float(d.norm() / ch.norm())})
        return err
    ef, er = errors(f32, f64), errors(r32, r64)
    line = {
        "metric": "FP32 intermediates vs FP64: max relative error after 50 RK4 steps",
        "value": max(e["max_rel_state"] for e in ef), "unit": "relative", "n_gpus": 1, "steps": nsteps,
        "warmup": 1, "ms_per_step": round(1e3 * t_f32, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 arithmetic, f32 storage of the viscous intermediates",
        "data": "synthetic (seeded acoustic pulse on uniform Mach 0.5 flow, periodic faces)",
        "config": {"workload": "viscous RHS, 512x512x256, periodic, 50 RK4 steps with GV10 filter", "dt": dt,
                   "fused_ms_per_step_fp64": round(1e3 * t_f64, 3),
                   "fused_ms_per_step_fp32_intermediates": round(1e3 * t_f32, 3),
                   "fused_speedup": round(t_f64 / t_f32, 4),
                   "fused_fp32_errors": ef,
                   "per_operator_ms_per_step_fp64": round(1e3 * t_r64, 3),
                   "per_operator_ms_per_step_fp32_solves_one_refinement": round(1e3 * t_r32, 3),
                   "per_operator_fp32_solves_max_rel_state": max(e["max_rel_state"] for e in er),
                   "per_operator_fp32_solve_errors": er},
    }
    print(json.dumps(line), flush=True)


def adjoint_bench(args, n=128, nsteps=40, every=8, plane=96, ndir=32):
    """Config 5: Euler at 128^3, periodic x, non-reflective y and z; the
    initial state is uniform Mach 0.5 flow plus an isentropic density
    perturbation rho' (seeded Gaussian, amplitude 1e-4, p' = c0^2 rho');
    40 RK4 steps with the filter and dt fixed from the initial CFL limit
    (CFL 0.5); loss = acoustic energy sum(p'^2/(2 rho0 c0^2) + rho0 |u'|^2/2)
    on the plane z = 96.  The gradient with respect to rho' comes from the
    adjoint with checkpoints every 8 steps (each segment re-run recording
    step tapes, then stepped backwards; compact_schemes_amd.rollout).
    Central finite differences on 32 random directions (N(0,1) white noise,
    numpy seeds 500..531) check it direction by direction, with these faces
    and with periodic faces."""
    import numpy as np
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.rollout import AcousticRollout, checkpointed_vjp
    NBC5 = [[45, 10, 10], [45, 10, 10]]
    curv = bool(getattr(args, "curvilinear", False))
    ro = AcousticRollout((n, n, n), NBC5, nsteps=nsteps, every=every, plane=plane, curvilinear=curv)
    sv = ro.solver
    q0 = ro.initial(ro.rho1)
    tapes = [sv.new_tape() for _ in range(every)]

    def forward(keep):
        q = q0.clone()
        ck = {}
        for k in range(nsteps):
            if keep and k % every == 0:
                ck[k] = q.clone()
            sv.rk4_step(q, dt=ro.dt)
        return q, ck

    def adjoint(ck, lam):
        for k0 in sorted(ck, reverse=True):
            x = ck[k0].clone()
            m = min(k0 + every, nsteps) - k0
            for i in range(m):
                sv.rk4_step_tape(x, tapes[i], dt=ro.dt)
            for i in reversed(range(m)):
                sv.rk4_vjp_tape(tapes[i], lam, ro.dt)
        return lam

    def adjoint_recompute(ck, lam):
        # the same with cs_solver_rk4_vjp, which recomputes the stages of each step
        for k0 in sorted(ck, reverse=True):
            states = [ck[k0]]
            for k in range(k0, min(k0 + every, nsteps) - 1):
                x = states[-1].clone()
                sv.rk4_step(x, dt=ro.dt)
                states.append(x)
            for x in reversed(states):
                sv.rk4_vjp(x, lam, ro.dt)
        return lam

    def timed(fn, reps=3):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            dt_ = time.perf_counter() - t0
            best = dt_ if best is None else min(best, dt_)
        return best, out

    forward(False)                                          # warm-up
    t_fwd, (qT, ck) = timed(lambda: forward(True))
    L0, lam0 = ro.loss_and_grad(qT)
    t_adj, lam = timed(lambda: adjoint(ck, lam0.clone()))
    t_rec, lam_rec = timed(lambda: adjoint_recompute(ck, lam0.clone()), reps=1)
    agree = float((lam - lam_rec).abs().max() / lam_rec.abs().max())
    L1, lam1 = checkpointed_vjp(sv, q0, nsteps, every, ro.dt, ro.loss_and_grad, tapes)
    assert torch.equal(lam1, lam)

    def fd_summary(rollout, eps):
        _, grad = rollout.gradient()
        fd, ad = rollout.fd_check([rollout.random_direction(500 + d) for d in range(ndir)], eps=eps, grad=grad,
                                  order=4)
        rel = np.abs(fd - ad) / np.abs(ad)
        return {"steps": nsteps, "directions": f"{ndir} x N(0,1) white noise, seeds 500..{500 + ndir - 1}",
                "stencil": "fourth-order central differences", "eps": eps, "max_rel_err": float(rel.max()),
                "median_rel_err": float(np.median(rel)),
                "normwise_rel_err": float(np.max(np.abs(fd - ad)) / np.max(np.abs(ad)))}
    fd_full = fd_summary(ro, 1e-5)
    fd_per = fd_summary(AcousticRollout((n, n, n), wl.NBC_PERIODIC, nsteps=nsteps, every=every, plane=plane,
                                        curvilinear=curv), 1e-5)
    line = {
        "metric": "adjoint/forward cost ratio, 40 Euler RK4 steps at 128^3 (checkpoints every 8)",
        "value": round(t_adj / t_fwd, 3), "unit": "ratio", "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": round(1e3 * t_fwd / nsteps, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (uniform Mach 0.5 flow + seeded isentropic density pulse, amplitude 1e-4)",
        "config": {"workload": "Euler 128^3, periodic x, non-reflective y/z, GV10 filter, 40 RK4 steps, loss = "
                               "acoustic energy on the plane z=96, gradient w.r.t. the initial density perturbation"
                               + (", sine-skewed curvilinear grid (per-point metrics)" if curv else ""),
                   "loss": L0, "dt": ro.dt, "forward_s": round(t_fwd, 4), "adjoint_s": round(t_adj, 4),
                   "adjoint": "per segment: the steps re-run with cs_solver_rk4_step_tape, then "
                              "cs_solver_rk4_vjp_tape backwards (fused transposed sweeps); best of 3",
                   "adjoint_s_stage_recompute_variant": round(t_rec, 4),
                   "ratio_stage_recompute_variant": round(t_rec / t_fwd, 3),
                   "tape_vs_recompute_max_rel": agree,
                   "fd_check_these_faces_40_steps": fd_full,
                   "fd_check_periodic_faces_40_steps": fd_per,
                   "checkpoint_fields": 5 * len(ck), "tape_doubles_per_step": sv.tape_size()},
    }
    print(json.dumps(line), flush=True)


def fp64_flops_from_profile(kernel, n):
    """FP64 flops per launch of `kernel` from the committed rocprofv3 --pmc
    summary of the FP64 instruction counters (64 lanes per wave instruction,
    FMA counted twice); None when the summary does not cover it."""
    if n != 256 or not os.path.exists(PROFILE_FP64):
        return None
    for line in open(PROFILE_FP64):
        parts = line.strip().split("\t")
        if len(parts) == 3 and parts[0] == "fp64_flops_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def source_digest():
    """sha256 of the kernel sources (compact_schemes_amd/csrc/*.hip, *.hpp and
    include/compact_schemes.h, in name order): the committed counter summaries
    record the digest of the sources they were measured with."""
    import glob
    import hashlib
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(ROOT, "compact_schemes_amd", "csrc", "*.hip")) +
                   glob.glob(os.path.join(ROOT, "compact_schemes_amd", "csrc", "*.hpp")))
    files.append(os.path.join(ROOT, "include", "compact_schemes.h"))
    for f in files:
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def profile_digest(path):
    """the source digest a counter summary records, or None"""
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.strip().split("\t")
        if len(parts) == 2 and parts[0] == "source_sha256":
            return parts[1]
    return None


def traffic_from_profile(kernel, n):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the plain-copy factor);
    None when the summary does not cover this kernel at this size."""
    if n != 256 or not os.path.exists(PROFILE_TRAFFIC):
        return None
    for line in open(PROFILE_TRAFFIC):
        parts = line.strip().split("\t")
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def adjoint_probe(args, n=128):
    """--reps RK4 steps of the config-5 solver recording tapes, then their
    adjoints backwards, twice; for kernel traces of the adjoint sweeps."""
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    st = wl.flow_state((n, n, n), seed=5, amp=1e-4)
    sv = FusedSolver((n, n, n), [[45, 10, 10], [45, 10, 10]], st, viscous=False, cfl=0.5)
    q = torch.from_numpy(st["qa"]).cuda()
    dt = sv.time_step(q)
    tapes = [sv.new_tape() for _ in range(args.reps)]
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    lam = torch.randn(q.numel(), generator=g, dtype=torch.float64, device="cuda")
    for _ in range(2):
        x = q.clone()
        for i in range(args.reps):
            sv.rk4_step(x.clone(), dt=dt)
        for i in range(args.reps):
            sv.rk4_step_tape(x, tapes[i], dt=dt)
        for i in reversed(range(args.reps)):
            sv.rk4_vjp_tape(tapes[i], lam, dt)
    torch.cuda.synchronize()
    print(json.dumps({"adjoint_probe": "done", "steps": args.reps, "n": n}), flush=True)


def batch_probe(args):
    """--reps steps (derivative x, y, z, then filter x, y, z) of the
    96-subdomain batch in grid mode 2, the mode the autotuner picks on MI355X
    (one launch of the single-subdomain kernels per subdomain and operator),
    through cs_batch_deriv and cs_batch_filter on one stream (the same
    launches cs_batch_step deals over eight forked streams; counter
    collection serialises dispatches, so the probe keeps one stream), and a
    plain copy of known size; per-dispatch counters are summed per kernel."""
    import torch
    from compact_schemes_amd import batch as bm
    from compact_schemes_amd.workloads import batch_fields
    per = [[45, 45, 45], [45, 45, 45]]
    dims, f0 = batch_fields()
    f = [t.clone() for t in f0]
    d = [torch.empty_like(f[s // 3]) for s in range(3 * len(dims))]
    bm.tune_clear()
    b = bm.Batch(dims, per)
    b.bind(f, d)
    b.set_persistent(2)
    st = torch.cuda.current_stream()
    for _ in range(args.reps):
        b.deriv(st)
        b.filter(st)
    npts = sum(x[0] * x[1] * x[2] for x in dims)
    src = torch.ones(npts, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"batch_probe": "done", "steps": args.reps, "subdomains": len(dims), "points": npts,
                      "copy_bytes_read": 8 * npts, "copy_bytes_written": 8 * npts}), flush=True)


def pmc_probe(args):
    """The sweeps of one RK stage and a plain copy of known size, --reps
    launches each, so that per-dispatch counters are easy to attribute."""
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FILTER_SWEEPS, FusedSolver, RK_STAGE_SWEEPS, sweep_names
    n = args.n
    dims = (n, n, n)
    lmx = n ** 3
    st = wl.flow_state(dims, seed=2)
    q = torch.from_numpy(st["qa"]).cuda()
    solver = FusedSolver(dims, wl.NBC_NON_REFLECTIVE, st)
    stream = torch.cuda.current_stream()
    solver.time_step(q, stream=stream)
    for w in RK_STAGE_SWEEPS:                 # the inputs of every sweep
        solver.sweep(w, q, stream=stream)
    for w in RK_STAGE_SWEEPS:
        for _ in range(args.reps):
            solver.sweep(w, q, stream=stream)
    solver.de.copy_(q)
    for w in FILTER_SWEEPS:                   # the filter sweeps, in place on the scratch field
        for _ in range(args.reps):
            solver.sweep(w, solver.de, solver.de, stream=stream)
    src = torch.ones(5 * lmx, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    names = sweep_names(solver.order)
    print(json.dumps({"sweeps": {names[w]: solver.kernel_name(w) for w in RK_STAGE_SWEEPS}, "reps": args.reps,
                      "copy_bytes_read": 8 * 5 * lmx, "copy_bytes_written": 8 * 5 * lmx}), flush=True)


def cpu_baseline(st, dims, reps):
    """The CPU oracle's full RHS (OpenMP over lines, all host cores) on the
    same workload; median of `reps` RHS evaluations after one warm-up."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.lib()
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    # the host share of one GPU: OMP_NUM_THREADS where the launcher sets it
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    pyoracle.set_num_threads(cores)
    lmx = dims[0] * dims[1] * dims[2]
    ctx = pyoracle.Context(dims, [[10, 10, 10], [10, 10, 10]])
    de, p = np.zeros(5 * lmx), np.zeros(lmx)
    ts = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        ctx.rhs(1, st["qa"], de, p, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"], st["sbcc"],
                st["umf"])
        if r > 0:
            ts.append(time.perf_counter() - t0)
    t = statistics.median(ts)
    return {"value": round(lmx / t / 1e9, 5), "unit": "Gpts/s", "cores": pyoracle.num_threads(), "kind": "port",
            "sample": f"{reps} full viscous RHS evaluations at {dims[0]}^3 (median, after 1 warm-up)",
            "cpu": cpu_model(), "s_per_rhs": round(t, 3)}


if __name__ == "__main__":
    main()
