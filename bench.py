"""Benchmark: DOF-matvec/s of the matrix-free hexa8 PCG solve on one MI355X.

Workload (the headline line): 96 x 48 x 48 hexa8 cantilever (221,184 elements,
691,488 free DOFs), E0 = 1, Emin = 1e-9, nu = 0.3, p = 3, densities i.i.d.
U[0.01, 1] (seed 20261203), x = 0 face clamped, unit load on the edge
{x = L, z = 0}; one solve to relative residual 1e-8.  A step is one such solve in
the mixed-precision form (FP32 operator, FP64 residual replacement), replayed from
captured HIP graphs as a caller would run it; the FP64 plain form (the Eigen loop)
and the fused FP64 form are timed beside it, and the 192 x 96 x 96 cantilever
(1,769,472 elements) is measured the same way under "config.large".

value = free DOFs x (operator applications of the step) / step wall time.

--pmc-workload runs a bounded counter workload instead (a 256 MiB device copy of
known size, then eager mixed-form iterations), for rocprofv3 --pmc passes.
--batched measures 64 independent jobs (sizes and loads drawn from seed 20261204:
Q4 60x20 .. 240x80 and hexa8 32x16x16 .. 96x48x48) solved one after another and as
one batch (icx_solve_batch), and checks that every displacement is bit-identical.
--serve runs the serving loop (icx.serve): Poisson arrivals, footprint admission,
OC design iterations with batched solves; p50/p99 design-iteration latency.
--design-loop runs the 192 x 96 x 96 cantilever design loop: Helmholtz filter R = 3
elements, 20 OC iterations from x0 = 0.3, mixed-precision solves; then, on the final
design, the error of u, compliance and dc/dx against the FP64 plain form at rel-res 1e-12.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "icarat-contact_amd"))

HBM_PEAK = 8.0e12
SEED = 20261203
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_hex8_fusedA.md")
TRAFFIC_FILE_SR = os.path.join(ROOT, "profiles", "hbm_traffic_hex8_sr.md")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, nargs=3, default=[96, 48, 48])
    ap.add_argument("--large", type=int, nargs=3, default=[192, 96, 96])
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel eagerly")
    ap.add_argument("--no-pipeline", action="store_true", help="poll the solver state chunk by chunk")
    ap.add_argument("--single-reduction", action="store_true",
                    help="mixed form with one kernel per iteration instead of kernels A and B")
    ap.add_argument("--pmc-workload", action="store_true")
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--design-loop", action="store_true")
    ap.add_argument("--loop-iters", type=int, default=20)
    ap.add_argument("--serve", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=60)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-iters", type=int, default=10)
    ap.add_argument("--serve-elems", type=float, nargs=2, default=[5e3, 3e5])
    ap.add_argument("--serve-budget", type=float, default=240e9)
    ap.add_argument("--serve-limit", type=int, default=None, help="serve the first N jobs of the stream")
    ap.add_argument("--serve-offset", type=int, default=0,
                    help="start at job N of the stream (with --serve-limit: one segment of it)")
    ap.add_argument("--serve-dump", default=None, help="write the latencies of the run to this .npz file")
    ap.add_argument("--serve-combine", nargs="+", default=None, metavar="STEM",
                    help="combine segments of one stream: STEM.json (the segment's serve line) and STEM.npz "
                         "(its --serve-dump) for each segment; no GPU")
    ap.add_argument("--serve-max-dofs", type=float, default=None,
                    help="admit jobs while the resident DOFs (of a size class) stay within this many")
    ap.add_argument("--serve-classes", type=float, nargs="*", default=None,
                    help="DOF bounds of the size classes, each served on a stream and batch of its own")
    ap.add_argument("--pmc-iters", type=int, default=200)
    ap.add_argument("--output", action="store_true", help="time the output step (mises, reactions, VTK)")
    ap.add_argument("--heat", action="store_true", help="heat-conduction solve (FP64 plain form)")
    ap.add_argument("--mfma", action="store_true", help="FP32 hexa8 matvec: Haar VALU kernel vs dense-K0 MFMA kernel")
    ap.add_argument("--mdpa", action="store_true", help="unstructured hexa8 mesh from an MDPA file (FP64 plain form)")
    ap.add_argument("--heat-div", type=int, default=100)
    ap.add_argument("--gcmma", action="store_true", help="GCMMA design loop (driver 2) on the C1 cantilever")
    ap.add_argument("--c1", action="store_true", help="the C1 design loop: OC driver, Q4 120 x 40, 50 iterations")
    ap.add_argument("--stress", action="store_true",
                    help="stress-constrained driver (C++ executable) on the reference's 200 x 100 job")
    ap.add_argument("--opt-steps", type=int, default=30)
    ap.add_argument("--q4", type=int, nargs=2, default=None, metavar=("NX", "NY"),
                    help="single solves of a quad4 cantilever NX x NY (seeded densities): FP64 fused and mixed")
    ap.add_argument("--q4-maxit", type=int, default=200000, help="iteration limit of the --q4 solves")
    return ap.parse_args()


def cpu_baseline(nx, ny, nz, seconds):
    """The oracle's CSR Jacobi-PCG (Eigen loop) on the host cores, bounded sample."""
    import ctypes as C

    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    pr = O.cantilever_3d(nx, ny, nz)
    rho = O.random_density(nx * ny * nz, SEED)
    t0 = time.perf_counter()
    b = pr.assemble(rho)
    t_asm = time.perf_counter() - t0
    # time PCG passes with a growing iteration cap until the sample reaches the budget
    its, dt = 0, 0.0
    cap = 20
    while True:
        x = np.zeros(pr.numeq)
        err = C.c_double()
        t0 = time.perf_counter()
        it = O.lib().icxo_sys_cg(pr.system(), b, x, 1e-8, cap, C.byref(err))
        dt = time.perf_counter() - t0
        its = it + 1 if err.value < 1e-8 else it
        if dt > seconds / 3 or err.value < 1e-8:
            break
        cap = int(cap * max(2.0, (seconds / 2) / max(dt, 1e-3)))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    return {
        "value": pr.numeq * its / dt, "unit": "DOF-matvec/s", "cores": threads,
        "kind": "port",
        "sample": "oracle CSR Jacobi-PCG (Eigen ConjugateGradient loop), %d iterations on the "
                  "same %dx%dx%d system (assembly %.2f s not included)" % (its, nx, ny, nz, t_asm),
        "assembly_s": t_asm, "pcg_s": dt, "iterations": its,
    }


def measured_traffic(nx, ny, nz, sr=False):
    """HBM bytes per launch of the iteration kernel from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known size)."""
    try:
        text = open(TRAFFIC_FILE_SR if sr else TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"traffic\[%dx%dx%d\]\s*=\s*([0-9.eE+]+)" % (nx, ny, nz), text)
    return float(m.group(1)) if m else None


def pmc_workload(args):
    import torch

    import icx
    a = torch.ones(64 * 1024 * 1024, device="cuda")  # 256 MiB
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy: 268435456 bytes read and written per launch", flush=True)
    nx, ny, nz = args.size
    ctx = icx.Context(0)
    ctx.set_option(icx.OPT_GRAPHS, 0)
    job = icx.cantilever_3d(ctx, nx, ny, nz)
    job.set_design(icx.random_density(job.nelx, SEED))
    _, info = job.solve(icx.SOLVE_MIXED, tol=args.tol, maxit=args.pmc_iters, want_u=False)
    ctx.synchronize()
    print("mixed form: %d iterations of %dx%dx%d (N = %d, Ne = %d)" % (
        info["iterations"], nx, ny, nz, job.numeq, job.nelx), flush=True)


def batched(args):
    """64 jobs solved one after another, then as one batch."""
    import numpy as np

    import icx
    ctx = icx.Context(0)
    if args.single_reduction:
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 1)
    rng = np.random.default_rng(np.random.PCG64(20261204))
    jobs = [icx.random_job(ctx, rng) for _ in range(args.batch_jobs)]
    for j in jobs:  # warm-up: buffers, captured graphs
        j.solve(icx.SOLVE_MIXED, tol=args.tol, maxit=50000, want_u=False)
    ctx.solve_batch(jobs, tol=args.tol)
    ctx.synchronize()
    t0 = time.perf_counter()
    seq = []
    for j in jobs:
        seq.append(j.solve(icx.SOLVE_MIXED, tol=args.tol, maxit=50000, want_u=True))
    ctx.synchronize()
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    infos = ctx.solve_batch(jobs, tol=args.tol)
    ctx.synchronize()
    t_bat = time.perf_counter() - t0
    same = all(np.array_equal(j.displacement(), u) and inf["iterations"] == si["iterations"]
               for j, (u, si), inf in zip(jobs, seq, infos))
    work = sum(j.numeq * inf["matvecs"] for j, inf in zip(jobs, infos))
    n2 = sum(1 for j in jobs if j.div[2] == 1)
    line = {"metric": "DOF-matvec/s, %d batched jobs, FP32-refined PCG to rel-res %g" % (len(jobs), args.tol),
            "value": work / t_bat, "unit": "DOF-matvec/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32 operator / f64 residual", "data": "synthetic (seeded sizes, loads, densities)",
            "config": {"workload": "batched jobs: %d Q4 (60x20..240x80) + %d hexa8 (32x16x16..96x48x48), seed "
                                   "20261204" % (n2, len(jobs) - n2),
                       "jobs": len(jobs), "free_dofs": sum(j.numeq for j in jobs),
                       "batched_s": t_bat, "sequential_s": t_seq, "sequential_dof_matvec_per_s": work / t_seq,
                       "speedup_batched_vs_sequential": t_seq / t_bat,
                       "bit_identical_to_sequential": same,
                       "iterations_min_max": [min(i["iterations"] for i in infos),
                                              max(i["iterations"] for i in infos)],
                       "all_converged": all(i["relres"] < args.tol for i in infos)}}
    print(json.dumps(line))


def output_step(args):
    """The output step on the large mesh after one mixed solve: element-mean von Mises
    stress and reactions (device work + copy of the result), and the binary VTK file
    (queued device work and pinned copies, then the file write)."""
    import tempfile

    import numpy as np

    import icx
    nx, ny, nz = args.large
    ctx = icx.Context(0)
    job = icx.cantilever_3d(ctx, nx, ny, nz)
    job.set_design(icx.random_density(job.nelx, SEED))
    job.solve(icx.SOLVE_MIXED, tol=args.tol, maxit=50000, want_u=False)
    ctx.synchronize()

    def timed(fn, reps=5):
        fn()
        ctx.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    mises_ms = timed(job.mises)
    react_ms = timed(job.reactions)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "res1.vtk")
        begin_ms = timed(job.output_begin, reps=1)
        job.output_write(path)
        job.output_begin()
        ctx.synchronize()
        t0 = time.perf_counter()
        job.output_write(path)
        write_ms = (time.perf_counter() - t0) * 1e3
        size = os.path.getsize(path)
    nn, ne = job.numnp, job.nelx
    line = {"metric": "output step: von Mises, reactions, binary VTK", "value": (begin_ms + write_ms),
            "unit": "ms per design iteration", "n_gpus": 1, "higher_is_better": False,
            "dtype": "f64 (file: f32)", "data": "synthetic (seeded densities)",
            "config": {"workload": "hexa8 cantilever %dx%dx%d after one mixed solve" % (nx, ny, nz),
                       "nodes": nn, "elements": ne,
                       "mises_ms_incl_copy": mises_ms, "reactions_ms_incl_copy": react_ms,
                       "vtk_device_and_copies_ms": begin_ms, "vtk_file_write_ms": write_ms,
                       "vtk_file_MB": size / 1e6, "vtk_write_MBps": size / 1e6 / (write_ms / 1e3)}}
    print(json.dumps(line))


def _hex8_matvec_blocks(nx, ny, nz, depth=0):
    """Blocks and element layers per block of the MODE 0 hexa8 operator tiling
    (15 x 15 node tiles, node layers along z, one element layer below each brick)."""
    sx, sy, sz = nx + 1, ny + 1, nz + 1
    cols = -(-sx // 15) * -(-sy // 15)

    def blocks(t):
        return cols * -(-sz // t)
    tk = 16
    if depth:
        tk = depth
    else:
        while tk > 2 and blocks(tk) < 2048:
            tk = tk >> 1 if tk > 4 else tk - 1
    while blocks(tk) > 4096:
        tk <<= 1
    tz = -(-sz // tk)
    layers = sum(min(k0 + tk, sz) - (k0 - 1) for k0 in range(0, sz, tk))
    return cols * tz, cols * layers


def mfma_matvec(args):
    """The FP32 hexa8 matvec y = K(E) x two ways on the same job: the Haar-block kernel on
    the vector units (the operator of the solve forms) and the dense-K0 contraction on the
    matrix cores (v_mfma_f32_32x32x2f32, ICX_OPT_MFMA_MATVEC), each timed per launch with
    launch-bound events, at the headline and the large cantilever."""
    import numpy as np

    import icx
    ctx = icx.Context(0)
    out = {}
    reps = max(args.steps, 5)
    for name, (nx, ny, nz) in (("headline", args.size), ("large", args.large)):
        job = icx.cantilever_3d(ctx, nx, ny, nz)
        job.set_design(icx.random_density(job.nelx, SEED))
        xv = np.random.default_rng(SEED).standard_normal(job.numeq)
        res = {}
        ys = {}
        for form, opt in (("haar_valu", 0), ("dense_mfma", 1)):
            ctx.set_option(icx.OPT_MFMA_MATVEC, opt)
            ys[form] = job.matvec(xv, prec=1)
            ctx.profile_begin(icx.PROF_MATVEC32)
            for _ in range(reps):
                job.matvec(xv, prec=1)
            sec, nl = ctx.profile_end()
            res[form] = {"us": sec / nl * 1e6, "launches": nl}
        ctx.set_option(icx.OPT_MFMA_MATVEC, 0)
        nn, ne = job.numnp, job.nelx
        nb, elayers = _hex8_matvec_blocks(nx, ny, nz)
        applied = elayers * 256  # element slots of the launch (patch halo and repeated layer included)
        byts = (6 * nn + ne) * 4 + nn
        for form, r in res.items():
            r["GBps_algorithmic"] = byts / (r["us"] * 1e-6) / 1e9
            r["hbm_frac"] = r["GBps_algorithmic"] / 8000.0
        m = res["dense_mfma"]
        m["mfma_flops_issued"] = applied * 32 * 24 * 2
        m["mfma_TFLOPs_issued"] = m["mfma_flops_issued"] / (m["us"] * 1e-6) / 1e12
        m["mfma_frac_of_fp32_peak"] = m["mfma_TFLOPs_issued"] / 157.3
        m["dense_flops_per_element"] = 24 * 24 * 2
        res["haar_valu"]["flops_per_element"] = 240
        d = np.abs(ys["dense_mfma"] - ys["haar_valu"]).max() / np.abs(ys["haar_valu"]).max()
        out[name] = {"workload": "hexa8 cantilever %dx%dx%d, FP32 y = K(E) x" % (nx, ny, nz), "elements": ne,
                     "nodes": nn, "blocks": nb, "bytes_per_launch": byts, "max_rel_diff_mfma_vs_valu": d, **res}
        job.close()
    ctx.close()
    line = {"metric": "FP32 hexa8 matvec per launch: Haar VALU vs dense-K0 MFMA", "n_gpus": 1, "dtype": "f32",
            "data": "synthetic (seeded densities)", "config": out}
    print(json.dumps(line))


def heat_solve(args):
    """Heat conduction on a hexa8 box (the 3D example of opt-heat: unit cube, all faces
    at T = 0, uniform heat source, conductivities 100 / 0.1, design 0.25, p = 3), one
    FP64 plain-form solve to 1e-8 on the random design of the other workloads; the
    operator is timed per launch with launch-bound events."""
    import numpy as np

    import icx
    n = args.heat_div
    ctx = icx.Context(0)
    if args.no_graphs:
        ctx.set_option(icx.OPT_GRAPHS, 0)
    job = icx.Job(ctx, icx.HEXA8, (n, n, n), (1.0, 1.0, 1.0), icx.HEAT)
    for face in ([0, 1, 0, 1, 0, 0], [0, 1, 0, 1, 1, 1], [0, 0, 0, 1, 0, 1], [1, 1, 0, 1, 0, 1],
                 [0, 1, 0, 0, 0, 1], [0, 1, 1, 1, 0, 1]):
        job.add_dirichlet([1, 0, 0, 0, 0, 0], [0] * 6, face)
    job.make_dof()
    job.add_neumann([0.1, 0, 0], [0, 1, 0, 1, 0, 1])
    job.set_conductivity((100.0, 100.0, 100.0, 0, 0, 0), (0.1, 0.1, 0.1, 0, 0, 0))
    job.set_design(0.05 + 0.95 * icx.random_density(job.nelx, SEED))
    job.solve(icx.SOLVE_PLAIN, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, info = job.solve(icx.SOLVE_PLAIN, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.profile_begin(icx.PROF_MATVEC64)
    _, pi = job.solve(icx.SOLVE_PLAIN, tol=args.tol, maxit=200000, want_u=False)
    s, nl = ctx.profile_end()
    mv_us = s / max(nl, 1) * 1e6
    N, nn, ne = job.numeq, job.numnp, job.nelx
    bytes_mv = 2 * nn * 8 + ne * 8 + nn  # x read, y write, weights, mask
    line = {"metric": "heat conduction: FP64 Jacobi-PCG (Eigen loop) DOF-matvec/s", "value": N * info["matvecs"] / wall,
            "unit": "DOF-matvec/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "heat hexa8 unit cube %d^3, faces at T = 0, source 0.1, k 100 / 0.1, p = 3"
                                   % n, "free_dofs": N, "iterations": info["iterations"], "relres": info["relres"],
                       "ms_per_solve": wall * 1e3, "operator_us": mv_us,
                       "operator_GBps_algorithmic": bytes_mv / (mv_us * 1e-6) / 1e9,
                       "operator_bytes_per_launch": bytes_mv}}
    print(json.dumps(line))


def gcmma_loop(args):
    """The GCMMA driver (opt_macro_gcmma.cpp:83-216) as the device-resident design loop on
    the C1 cantilever (Q4 120 x 40, plane stress, point load, cone filter r = 2 h, x0 = 0.5),
    FP64 fused-form solves (the Eigen CG loop's algorithm) to 1e-10 (the reference's
    Eigen CG tolerance).  Every solve of a
    step (the design's and each trial design's) is in the timed region.  CPU baseline: the
    oracle's loop with its CG at the same tolerance, same steps."""
    import numpy as np

    import icx
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    steps = args.opt_steps
    ctx = icx.Context(0)
    if args.no_graphs:
        ctx.set_option(icx.OPT_GRAPHS, 0)
    job = icx.cantilever_2d(ctx, 120, 40)
    x0 = np.full(job.nelx, 0.5)
    job.opt_init(x0, driver=2, filter=icx.FILTER_CONE, radius=2.0, form=icx.SOLVE_FUSED, tol=1e-10,
                 maxit=100000)
    ctx.synchronize()
    t0 = time.perf_counter()
    recs = [job.opt_step() for _ in range(steps)]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    solves = steps + sum(r["inner_loops"] for r in recs) + steps
    cg = sum(r["cg_iterations"] + r["trial_cg_iterations"] for r in recs)
    cpu = None
    if not args.no_cpu:
        pr = O.cantilever_2d(120, 40)
        t1 = time.perf_counter()
        ref = pr.run_driver(2, x0, steps, 2.0, projection=True, direct=False, cg_tol=1e-10, cg_maxit=100000)
        tc = time.perf_counter() - t1
        cpu = {"value": steps / tc, "unit": "design iterations/s", "cores": os.cpu_count() or 1, "kind": "port",
               "sample": "oracle GCMMA loop, %d steps on the same job (CG 1e-10)" % steps,
               "objective_last_rel_diff": abs(ref["obj"][-1] - recs[-1]["objective"]) / ref["obj"][-1]}
    line = {"metric": "GCMMA design iterations/s", "value": steps / wall, "unit": "design iterations/s",
            "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "GCMMA driver, Q4 120 x 40 cantilever, cone r = 2, %d steps" % steps,
                       "ms_per_step": wall / steps * 1e3, "solves": solves, "pcg_iterations": cg,
                       "inner_loops": [r["inner_loops"] for r in recs],
                       "objective_first_last": [recs[0]["objective"], recs[-1]["objective"]]},
            "cpu_baseline": cpu}
    print(json.dumps(line))


def c1_loop(args):
    """The reference's CPU-runnable case: the OC driver (opt_macro_oc.cpp:84-172) on the
    Q4 120 x 40 plane-stress cantilever (E0 = 1, Emin = 1e-9, nu = 0.3, p = 3, left edge
    clamped, unit point load at the right mid-edge), cone filter r = 1.5, the reference's
    projection schedule, x0 = 0.5, 50 iterations, FP64 fused-form solves to 1e-10 (the
    reference's Eigen CG tolerance), the OC bisection on the device.  CPU baseline: the
    oracle's loop (CG 1e-10) on the same job."""
    import numpy as np

    import icx
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    steps = 50
    ctx = icx.Context(0)
    if args.no_graphs:
        ctx.set_option(icx.OPT_GRAPHS, 0)
    job = icx.cantilever_2d(ctx, 120, 40)
    x0 = np.full(job.nelx, 0.5)
    job.opt_init(x0, driver=0, filter=icx.FILTER_CONE, radius=1.5, form=icx.SOLVE_FUSED, tol=1e-10,
                 maxit=100000)
    ctx.synchronize()
    t0 = time.perf_counter()
    recs = [job.opt_step() for _ in range(steps)]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    cpu = None
    if not args.no_cpu:
        pr = O.cantilever_2d(120, 40)
        t1 = time.perf_counter()
        ref = pr.run_driver(0, x0, steps, 1.5, projection=True, direct=False, cg_tol=1e-10, cg_maxit=100000)
        tc = time.perf_counter() - t1
        cpu = {"value": steps / tc, "unit": "design iterations/s", "cores": os.cpu_count() or 1, "kind": "port",
               "sample": "oracle OC loop, the same 50 iterations (CG 1e-10)",
               "objective_max_rel_diff": float(np.max(np.abs(ref["obj"] - np.array([r["objective"] for r in recs]))
                                                      / ref["obj"]))}
    line = {"metric": "C1 design iterations/s", "value": steps / wall, "unit": "design iterations/s",
            "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "OC driver, Q4 120 x 40 cantilever, cone r = 1.5, projection, 50 iterations",
                       "ms_per_step": wall / steps * 1e3,
                       "pcg_iterations": int(sum(r["cg_iterations"] for r in recs)),
                       "objective_first_last": [recs[0]["objective"], recs[-1]["objective"]]},
            "cpu_baseline": cpu}
    print(json.dumps(line))


STRESS_TOML = """[mesh]
dimension = 2
division = [200, 100, 1]
length = [200.0, 100.0, 0.25]
type = "quad4"

[boundary_condition]
dirichlet = [
  [[1, 1, 0, 0, 0, 0], [0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [[0.0, 0.0], [0.0, 100.0], [0.0, 0.0]]],
]
neumann = [
  [[0.0, -0.1, 0.0], [[200.0, 200.0], [40.0, 60.0], [0.0, 0.0]]],
]

[material]
designs_0 = 0.5
penalty = 3.0
poisson = 0.3
state = "plane_stress"
type = "Elastic"
young = [0.001, 100.0]

[this_problem]
fdm = false
p_norm = 8.0
qp = 0.7
stress_limit = 1.0
"""


def stress_driver(args):
    """The stress-constrained driver (constraint/opt-stress-constraint, its init.toml:
    Q4 200 x 100, plane stress) run as the C++ executable over libicx for --opt-steps
    steps (structural solve, stress p-norm, adjoint solve, filter, MMA, binary VTK per
    step; structural and adjoint solves in the FP64 fused form of the Eigen CG loop at
    1e-10); value = design iterations per second of the executable's whole run.  CPU
    baseline: the oracle's loop (CG 1e-10 for both solves) on the same job."""
    import subprocess
    import tempfile

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    steps = args.opt_steps
    exe = os.path.join(ROOT, "icarat-contact_amd", "_build", "opt_stress_constraint")
    with tempfile.TemporaryDirectory() as d:
        toml = os.path.join(d, "init.toml")
        open(toml, "w").write(STRESS_TOML)
        t0 = time.perf_counter()
        out = subprocess.run([exe, toml, "HIP_PCG_FUSED", str(steps)], cwd=d, capture_output=True, text=True,
                             timeout=600)
        wall = time.perf_counter() - t0
        if out.returncode != 0:
            raise RuntimeError(out.stderr[-2000:])
        rows = np.loadtxt(os.path.join(d, "function.csv"), delimiter=",", ndmin=2)
    cpu = None
    if not args.no_cpu:
        pr = O.Problem(4, (200, 100, 1), (200.0, 100.0, 0.25), O.PLANE_STRESS, young=(0.001, 100.0))
        pr.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 0, 100, 0, 0])
        pr.make_dof()
        pr.add_neumann([0.0, -0.1, 0.0], [200, 200, 40, 60, 0, 0])
        k = min(steps, 5)
        t1 = time.perf_counter()
        ref = pr.run_driver(3, np.full(pr.nelx, 0.5), k, 2.0, projection=True, direct=False, cg_tol=1e-10,
                            cg_maxit=100000, stress=(1.0, 8.0, 0.7))
        tc = time.perf_counter() - t1
        cpu = {"value": k / tc, "unit": "design iterations/s", "cores": os.cpu_count() or 1, "kind": "port",
               "sample": "oracle stress-constrained loop, first %d steps of the same job (CG 1e-10)" % k,
               "constraint_rel_diff_first_steps": float(np.max(np.abs(ref["vol"] - rows[:k, 2]) /
                                                              np.maximum(1e-12, np.abs(ref["vol"]))))}
    line = {"metric": "stress-constrained design iterations/s", "value": steps / wall,
            "unit": "design iterations/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "stress-constrained driver, Q4 200 x 100 plane stress (init.toml), %d steps"
                                   % steps, "ms_per_step": wall / steps * 1e3,
                       "volume_first_last": [rows[0, 1], rows[-1, 1]],
                       "constraint_first_last": [rows[0, 2], rows[-1, 2]]},
            "cpu_baseline": cpu}
    print(json.dumps(line))


def hexa8_box_mesh(nx, ny, nz, amplitude, seed):
    """A hexa8 box mesh (h = 1) as explicit arrays in the reference's local node order,
    interior nodes moved by up to amplitude * h in every direction (seeded)."""
    import numpy as np
    sx, sy, sz = nx + 1, ny + 1, nz + 1
    k, j, i = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    X = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    rng = np.random.default_rng(seed)
    for d, n in enumerate((nx, ny, nz)):
        inner = (X[:, d] > 0) & (X[:, d] < n)
        X[inner, d] += amplitude * rng.uniform(-1.0, 1.0, inner.sum())
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = (ex + sx * (ey + sy * ez)).ravel()
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    conn = np.stack([base + a + sx * (b + sy * c) for a, b, c in corner], axis=1).astype(np.int32)
    return X.ravel(), conn.ravel()


def mesh_solve(args):
    """An unstructured hexa8 mesh read from an MDPA file: the C2 box with every interior
    node moved by up to 0.2 h, clamped at x = 0, line load on the edge x = L, z = 0; one
    FP64 plain-form solve to 1e-8 on the seeded random design.  The element pass of the
    operator is timed per launch with launch-bound events."""
    import tempfile

    import numpy as np

    import icx
    nx, ny, nz = args.size
    x, conn = hexa8_box_mesh(nx, ny, nz, 0.2, SEED)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "box.mdpa")
        X = x.reshape(-1, 3)
        E = conn.reshape(-1, 8) + 1
        with open(path, "w") as f:
            f.write("Begin Nodes\n")
            f.writelines("%d %.17g %.17g %.17g\n" % (n + 1, p[0], p[1], p[2]) for n, p in enumerate(X))
            f.write("End Nodes\n\nBegin Elements SmallDisplacementElement3D8N\n")
            f.writelines("%d 0 %s\n" % (e + 1, " ".join(map(str, nodes))) for e, nodes in enumerate(E))
            f.write("End Elements\n")
        t0 = time.perf_counter()
        xr, cr = icx.read_mdpa(path, icx.HEXA8)
        t_read = time.perf_counter() - t0
    ctx = icx.Context(0)
    if args.no_graphs:
        ctx.set_option(icx.OPT_GRAPHS, 0)
    t0 = time.perf_counter()
    job = icx.Job.from_mesh(ctx, icx.HEXA8, xr, cr, icx.SOLID3D)
    t_create = time.perf_counter() - t0
    job.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0, 0, 0, ny, 0, nz])
    job.make_dof()
    # total F_z = -1 on the edge x = L, z = 0 (boundary nodes are not moved), end nodes
    # carrying half weight, as the structured C2 case
    f = np.zeros(job.numeq)
    dof = job.dof_map()
    for jj in range(ny + 1):
        d = dof[3 * (nx + (nx + 1) * jj) + 2]
        if d < job.numeq:
            f[d] += -(0.5 if jj in (0, ny) else 1.0) / ny
    job.set_fext(f)
    job.set_design(icx.random_density(job.nelx, SEED))
    job.solve(icx.SOLVE_PLAIN, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, info = job.solve(icx.SOLVE_PLAIN, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.profile_begin(icx.PROF_MATVEC64)
    _, pi = job.solve(icx.SOLVE_PLAIN, tol=args.tol, maxit=200000, want_u=False)
    s, nl = ctx.profile_end()
    print("mesh profile window: %.6f s over %d launches, %d matvecs" % (s, nl, pi["matvecs"]), file=sys.stderr,
          flush=True)
    el_us = s / max(nl, 1) * 1e6
    # the mixed form on the same job: FP32 Gauss-point element pass, FP64 replacement
    job.solve(icx.SOLVE_MIXED, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, mi = job.solve(icx.SOLVE_MIXED, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    mixed_wall = time.perf_counter() - t0
    ctx.profile_begin(icx.PROF_FUSED_A)
    job.solve(icx.SOLVE_MIXED, tol=args.tol, maxit=200000, want_u=False)
    s32, nl32 = ctx.profile_end()
    el32_us = s32 / max(nl32, 1) * 1e6
    # the FP64 fused form on the same job: the plain form's element pass and gather, with
    # the vector updates fused into a direction kernel and kernel B
    job.solve(icx.SOLVE_FUSED, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, fi = job.solve(icx.SOLVE_FUSED, tol=args.tol, maxit=200000, want_u=False)
    ctx.synchronize()
    fused_wall = time.perf_counter() - t0
    # FilterHelmholtz on the same mesh (R = 2 element widths, the reference's 1e-10): the
    # scalar operator integrated per element in the same way
    job.helmholtz_setup(2.0)
    rho = icx.random_density(job.nelx, SEED)
    job.helmholtz(rho, tol=1e-10)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, h_it = job.helmholtz(rho, tol=1e-10)
    h_ms = (time.perf_counter() - t0) * 1e3
    N, nn, ne = job.numeq, job.numnp, job.nelx
    # element pass (Gauss-point integration from node coordinates, nothing per element
    # stored): reads connectivity (8 ints) and E per element, coordinates and values once
    # per node, writes y_e (24 words) per element.  FP64 operations per element: 8 Gauss
    # points x (296 fused multiply-adds + 86 multiplies/adds + 8 of the divide sequence)
    # = 8 x 686, plus 24 for E y_e (instruction count of k_mesh_element<8,3>).
    bytes_el = ne * (8 * 4 + 8 + 24 * 8) + nn * 6 * 8
    flops_el = ne * (8 * 686 + 24)
    tf = flops_el / (el_us * 1e-6) / 1e12 if el_us else None
    line = {"metric": "unstructured hexa8 (MDPA): FP64 Jacobi-PCG (Eigen loop) DOF-matvec/s", "value": N * info["matvecs"] / wall,
            "unit": "DOF-matvec/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic (seeded node perturbation and densities)",
            "config": {"workload": "hexa8 MDPA mesh %dx%dx%d box, interior nodes moved by up to 0.2 h, "
                                   "x = 0 clamped, edge load F_z = -1, rel-res %g" % (nx, ny, nz, args.tol),
                       "elements": ne, "nodes": nn, "free_dofs": N, "iterations": info["iterations"],
                       "relres": info["relres"], "ms_per_solve": wall * 1e3,
                       "mdpa_read_s": t_read, "job_create_s": t_create,
                       "element_pass_us": el_us,
                       "element_pass_bytes_per_launch": bytes_el,
                       "element_pass_GBps_algorithmic": bytes_el / (el_us * 1e-6) / 1e9 if el_us else None,
                       "element_pass_flops_per_launch": flops_el,
                       "helmholtz_filter": {"R": 2.0, "tol": 1e-10, "iterations": h_it, "ms": h_ms},
                       "fused_fp64": {"ms_per_solve": fused_wall * 1e3, "iterations": fi["iterations"],
                                      "relres": fi["relres"], "dof_matvec_per_s": N * fi["matvecs"] / fused_wall,
                                      "speedup_vs_plain": wall / fused_wall},
                       "mixed": {"ms_per_solve": mixed_wall * 1e3, "iterations": mi["iterations"],
                                 "refinements": mi["refinements"], "relres": mi["relres"],
                                 "dof_matvec_per_s": N * mi["matvecs"] / mixed_wall,
                                 "element_pass_fp32_us": el32_us,
                                 "speedup_vs_plain": wall / mixed_wall}},
            "roofline": {"bound": "fp64", "achieved": tf, "peak": 78.6, "unit": "TFLOP/s",
                         "frac": tf / 78.6 if tf else None, "traffic": None,
                         "kernel": "k_mesh_element<8,3> (Gauss-point element pass)"}}
    print(json.dumps(line))


def design_loop(args):
    import numpy as np

    import icx
    nx, ny, nz = args.large
    ctx = icx.Context(0)
    job = icx.cantilever_3d(ctx, nx, ny, nz)
    job.opt_init(np.full(job.nelx, 0.3), driver=0, filter=icx.FILTER_HELMHOLTZ, radius=3.0,
                 projection=True, form=icx.SOLVE_MIXED, tol=args.tol, maxit=50000, helm_tol=1e-10)
    recs = []
    t0 = time.perf_counter()
    for _ in range(args.loop_iters):
        recs.append(job.opt_step())
        print("step %d c=%.10g cg=%d %.3f s" % (recs[-1]["step"], recs[-1]["objective"], recs[-1]["cg_iterations"],
                                               recs[-1]["seconds"]), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    # the final step's design (E field) and mixed-form state: u, c, dc
    u_m = job.displacement()
    c_m = job.compliance()
    dc_m, _ = job.sensitivity()
    _, info_p = job.solve(icx.SOLVE_PLAIN, tol=1e-12, maxit=200000)
    u_p = job.displacement()
    c_p = job.compliance()
    dc_p, _ = job.sensitivity()
    line = {"metric": "design loop: 20 OC iterations with Helmholtz filter, mixed precision; error vs FP64",
            "value": args.loop_iters / wall, "unit": "design iterations/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32 operator / f64 residual", "data": "synthetic",
            "config": {"workload": "hexa8 cantilever %dx%dx%d, Helmholtz R = 3 elements, OC, x0 = 0.3, rel-res %g"
                                   % (nx, ny, nz, args.tol),
                       "wall_s": wall, "objective": [r["objective"] for r in recs],
                       "cg_iterations": [r["cg_iterations"] for r in recs],
                       "oc_bisections": [r["update_steps"] for r in recs],
                       "step_s": [r["seconds"] for r in recs],
                       "reference": "FP64 plain form (Eigen loop) at rel-res 1e-12, %d iterations, same design"
                                    % info_p["iterations"],
                       "err_u_rel_l2": float(np.linalg.norm(u_m - u_p) / np.linalg.norm(u_p)),
                       "err_c_rel": float(abs(c_m - c_p) / abs(c_p)),
                       "err_dc_rel_max": float(np.max(np.abs(dc_m - dc_p)) / np.max(np.abs(dc_p)))}}
    print(json.dumps(line))


def serving(args):
    import icx
    from icx import serve
    ctx = icx.Context(0)
    out = serve.run(ctx, n_jobs=args.serve_jobs, rate=args.serve_rate, iters=args.serve_iters,
                    elems=tuple(args.serve_elems), budget=args.serve_budget, tol=args.tol,
                    log=lambda m: print(m, file=sys.stderr, flush=True), limit=args.serve_limit,
                    max_resident_dofs=int(args.serve_max_dofs) if args.serve_max_dofs else None,
                    offset=args.serve_offset, classes=args.serve_classes)
    out.pop("designs", None)
    import numpy as np
    lat = out.pop("latencies_ms")
    waits = out.pop("admission_wait_ms")
    recs = np.array(out.pop("iteration_records"), dtype=float).reshape(-1, 10)
    out["latency_by_dofs"] = latency_by_dofs(recs)  # where the tail comes from
    if args.serve_dump:  # every design iteration's latency and every admission wait, for
        # combining the segments of a stream served in parts
        np.savez(args.serve_dump, latencies_ms=np.array(lat), admission_wait_ms=np.array(waits), records=recs)
    line = {"metric": "serving: design-iteration latency p50/p99 and DOF-matvec/s",
            "value": out["dof_matvec_per_s"], "unit": "DOF-matvec/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "hexa8 jobs: f32 operator / f64 residual; quad4 jobs: f64",
            "data": "synthetic (Poisson arrivals, seed 20261205)",
            "config": out}
    print(json.dumps(line))


def quad4_solve(args):
    """Single solves of a quad4 cantilever (plane stress, seeded U[0.01, 1] densities) to
    the tolerance in the FP64 fused and the mixed form: wall time of graph-replayed solves
    and kernel A / B per launch from HIP events."""
    import numpy as np
    import icx
    nx, ny = args.q4
    ctx = icx.Context(0)
    if args.no_graphs:
        ctx.set_option(icx.OPT_GRAPHS, 0)
    job = icx.cantilever_2d(ctx, nx, ny)
    job.set_design(icx.random_density(job.nelx, SEED))
    N, Ne, nodes = job.numeq, job.nelx, (nx + 1) * (ny + 1)
    forms = {}
    for name, form, w in (("fp64_fused", icx.SOLVE_FUSED, 8), ("mixed", icx.SOLVE_MIXED, 4)):
        for _ in range(args.warmup):
            job.solve(form, tol=args.tol, maxit=args.q4_maxit, want_u=False)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, info = job.solve(form, tol=args.tol, maxit=args.q4_maxit, want_u=False)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        k = {}
        for kn, kid in (("A", icx.PROF_FUSED_A), ("B", icx.PROF_FUSED_B)):
            ctx.profile_begin(kid)
            _, pi = job.solve(form, tol=args.tol, maxit=args.q4_maxit, want_u=False)
            secs, _ = ctx.profile_end()
            k[kn] = secs / max(pi["iterations"], 1) * 1e6
        bytes_A = (6 * N + nodes + Ne) * w  # reads r, p_old, x, invs, E; writes p, Ap, x
        forms[name] = {"ms_per_solve": ms, "iterations": info["iterations"], "relres": info["relres"],
                       "converged": info["converged"], "matvecs": info["matvecs"],
                       "dof_matvec_per_s": N * info["matvecs"] / (ms * 1e-3),
                       "kernel_A_us": k["A"], "kernel_B_us": k["B"], "kernel_A_bytes": bytes_A,
                       "kernel_A_GBps": bytes_A / (k["A"] * 1e-6) / 1e9}
    line = {"metric": "quad4 single solve: time to rel-res %g, FP64 fused and mixed" % args.tol,
            "value": forms["fp64_fused"]["ms_per_solve"], "unit": "ms", "n_gpus": 1, "higher_is_better": False,
            "dtype": "f64 (fused) / f32 operator + f64 residual (mixed)", "data": "synthetic (seeded densities)",
            "config": {"workload": "quad4 plane-stress cantilever %dx%d" % (nx, ny), "free_dofs": N,
                       "elements": Ne, "forms": forms}}
    print(json.dumps(line))


def latency_by_dofs(recs):
    """Design-iteration latency and PCG iterations by job size (free DOFs), from serving
    records (job, step, DOFs, PCG iterations, replacements, converged, latency ms, ...)."""
    import numpy as np
    by_size = {}
    for lo, hi in ((0, 1e5), (1e5, 1e6), (1e6, 1e12)):
        sel = recs[(recs[:, 2] >= lo) & (recs[:, 2] < hi)]
        if len(sel):
            by_size["%g-%g" % (lo, hi)] = {
                "design_iterations": int(len(sel)), "ms_p50": float(np.percentile(sel[:, 6], 50)),
                "ms_p99": float(np.percentile(sel[:, 6], 99)),
                "pcg_iterations_p50": float(np.percentile(sel[:, 3], 50)),
                "pcg_iterations_max": int(sel[:, 3].max())}
    return by_size


def serve_combine(stems):
    """One serve line for a stream served in segments (bench.py --serve --serve-offset K
    --serve-limit L --serve-dump STEM.npz > STEM.json, one segment per sitting, each
    starting empty): wall times and DOF-matvecs add up, the percentiles are over every
    design iteration and admission of all segments."""
    import numpy as np
    lines = [json.loads(open(s + ".json").read().strip().splitlines()[-1]) for s in stems]
    dumps = [np.load(s + ".npz") for s in stems]
    wall = sum(l["config"]["wall_s"] for l in lines)
    dmv = sum(l["value"] * l["config"]["wall_s"] for l in lines)
    lat = np.concatenate([d["latencies_ms"] for d in dumps])
    waits = np.concatenate([d["admission_wait_ms"] for d in dumps])
    recs = np.concatenate([d["records"] for d in dumps])
    segs = [{"offset": l["config"]["offset"], "jobs": l["config"]["jobs"], "wall_s": l["config"]["wall_s"],
             "iteration_ms_p50": l["config"]["iteration_ms_p50"], "iteration_ms_p99": l["config"]["iteration_ms_p99"],
             "unconverged_solves": l["config"].get("unconverged_solves")} for l in lines]
    c0 = lines[0]["config"]
    line = {"metric": "serving: design-iteration latency p50/p99 and DOF-matvec/s",
            "value": dmv / wall, "unit": "DOF-matvec/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": lines[0]["dtype"], "data": lines[0]["data"],
            "config": {"workload": "served in %d segments, each starting empty" % len(lines),
                       "jobs": int(sum(s["jobs"] for s in segs)), "design_iterations": int(len(lat)),
                       "wall_s": wall, "iteration_ms_p50": float(np.percentile(lat, 50)),
                       "iteration_ms_p99": float(np.percentile(lat, 99)),
                       "admission_wait_ms_p50": float(np.percentile(waits, 50)),
                       "admission_wait_ms_p99": float(np.percentile(waits, 99)),
                       "unconverged_solves": int((recs[:, 5] == 0).sum()),
                       "latency_by_dofs": latency_by_dofs(recs),
                       "iterations_per_job": c0["iterations_per_job"], "elements": c0["elements"],
                       "arrival_rate_per_s": c0["arrival_rate_per_s"], "budget_bytes": c0["budget_bytes"],
                       "max_resident_dofs": c0["max_resident_dofs"], "tol": c0["tol"], "segments": segs}}
    print(json.dumps(line))


def measure(icx, ctx, size, tol, steps, warmup, plain_too, sr=False):
    """Mixed-form solves of one cantilever: graph-replayed steps, then one solve with
    HIP events around every launch of the iteration kernel(s): the single-reduction
    kernel (sr), else the fused kernels A and B."""
    import numpy as np
    nx, ny, nz = size
    job = icx.cantilever_3d(ctx, nx, ny, nz)
    job.set_design(icx.random_density(job.nelx, SEED))
    N, Ne = job.numeq, job.nelx
    for _ in range(warmup):
        job.solve(icx.SOLVE_MIXED, tol=tol, maxit=50000, want_u=False)
    ctx.synchronize()
    infos = []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, info = job.solve(icx.SOLVE_MIXED, tol=tol, maxit=50000, want_u=False)
        infos.append(info)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # Dominant kernel: HIP events around every launch of the iteration kernel (the
    # library's stream; launched one by one while the window is open).
    # Single-reduction kernel, algorithmic bytes per launch: reads r, w, s, invdiag, p,
    # x_lo (6 words per DOF) and E (1 word per element); writes r, w, s, p, x_lo (5 words
    # per DOF).  Kernel A of the two-kernel form: reads r, p_old, x_lo (3 words per DOF),
    # the inverse modulus sum of every node (1 word per node) and E (1 word per element);
    # writes p, Ap, x_lo (3 words per DOF).
    prof = {}
    for name, kid in (("A", icx.PROF_FUSED_A), ("B", icx.PROF_FUSED_B)):
        if sr and name == "B":
            prof[name] = (0.0, 0, 1)
            continue
        ctx.profile_begin(kid)
        _, pinfo = job.solve(icx.SOLVE_MIXED, tol=tol, maxit=50000, want_u=False)
        secs, launches = ctx.profile_end()
        prof[name] = (secs, launches, pinfo["iterations"])
    kern_s, launches, executed = prof["A"]
    nodes = (size[0] + 1) * (size[1] + 1) * (size[2] + 1)
    bytes_A = ((11 * N + Ne) if sr else (6 * N + nodes + Ne)) * 4
    res = {
        "N": N, "Ne": Ne, "wall": wall, "infos": infos,
        "ms_step": wall / steps * 1e3,
        "value": N * sum(i["matvecs"] for i in infos) / wall,
        "bytes_A": bytes_A,
        "avg_A": kern_s / max(executed, 1),
        "launches_A": launches, "executed_A": executed,
        "achieved_A": bytes_A * executed / kern_s if kern_s > 0 else 0.0,
        "avg_B": prof["B"][0] / max(prof["B"][2], 1),
    }
    inner = infos[-1]["iterations"]
    refin = infos[-1]["refinements"]
    # whole-solve algorithmic bytes: per inner iteration 11 N w + N_e w_E (single
    # reduction) or 9 N w + N_e w_E (two kernels, the Ap round trip not counted), w = 4;
    # 6 N 8 + N_e 8 per FP64 residual replacement
    res["bytes_solve"] = inner * ((11 if sr else 9) * N * 4 + Ne * 4) + refin * (6 * N * 8 + Ne * 8)
    res["solve_frac"] = res["bytes_solve"] / (res["ms_step"] * 1e-3) / HBM_PEAK
    res["p50"] = float(np.percentile([i["seconds"] * 1e3 for i in infos], 50))
    res["p99"] = float(np.percentile([i["seconds"] * 1e3 for i in infos], 99))
    if plain_too:
        ctx.synchronize()
        t1 = time.perf_counter()
        _, pinfo = job.solve(icx.SOLVE_PLAIN, tol=tol, maxit=50000, want_u=False)
        ctx.synchronize()
        pt = time.perf_counter() - t1
        t1 = time.perf_counter()
        _, finfo = job.solve(icx.SOLVE_FUSED, tol=tol, maxit=50000, want_u=False)
        ctx.synchronize()
        ft = time.perf_counter() - t1
        res["plain"] = {"ms": pt * 1e3, "iterations": pinfo["iterations"], "relres": pinfo["relres"],
                        "dof_matvec_per_s": N * pinfo["matvecs"] / pt,
                        "speedup_mixed_vs_plain": pt * 1e3 / res["ms_step"],
                        "fused_fp64_ms": ft * 1e3, "fused_fp64_iterations": finfo["iterations"]}
    else:
        res["plain"] = None
    del job
    return res


def summary(size, r, tol):
    info = r["infos"][-1]
    return {"workload": "hexa8 cantilever %dx%dx%d, single solve to rel-res %g, mixed precision"
                        % (tuple(size) + (tol,)),
            "elements": r["Ne"], "free_dofs": r["N"], "inner_iterations": info["iterations"],
            "fp64_replacements": info["refinements"], "matvecs": info["matvecs"],
            "true_relres": info["relres"], "ms_per_solve": r["ms_step"],
            "dof_matvec_per_s": r["value"],
            "step_ms_p50": r["p50"], "step_ms_p99": r["p99"],
            "solve_hbm_frac": r["solve_frac"],
            "kernel_A_us": r["avg_A"] * 1e6, "kernel_A_GBps": r["achieved_A"] / 1e9,
            "kernel_B_us": r["avg_B"] * 1e6,
            "fp64_plain": r["plain"]}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the single-solve path would shard by x-slabs with one halo exchange per "
              "iteration; not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_workload:
        pmc_workload(args)
        return
    if args.batched:
        batched(args)
        return
    if args.q4:
        quad4_solve(args)
        return
    if args.serve_combine:
        serve_combine(args.serve_combine)
        return
    if args.serve:
        serving(args)
        return
    if args.output:
        output_step(args)
        return
    if args.gcmma:
        return gcmma_loop(args)
    if args.c1:
        return c1_loop(args)
    if args.stress:
        return stress_driver(args)
    if args.heat:
        heat_solve(args)
        return
    if args.mfma:
        mfma_matvec(args)
        return
    if args.mdpa:
        mesh_solve(args)
        return
    if args.design_loop:
        design_loop(args)
        return

    import icx
    ctx = icx.Context(0)
    if args.no_graphs:
        ctx.set_option(icx.OPT_GRAPHS, 0)
    if args.no_pipeline:
        ctx.set_option(icx.OPT_PIPELINE, 0)
    sr = args.single_reduction
    if sr:
        ctx.set_option(icx.OPT_SINGLE_REDUCTION, 1)
    r = measure(icx, ctx, args.size, args.tol, args.steps, args.warmup, not args.no_plain, sr)
    print("headline %dx%dx%d: %.2f ms/solve" % (tuple(args.size) + (r["ms_step"],)), file=sys.stderr,
          flush=True)
    large = None
    if not args.no_large:
        rl = measure(icx, ctx, args.large, args.tol, 1, 1, not args.no_plain, sr)
        large = summary(args.large, rl, args.tol)
        large["roofline_A"] = {"achieved": rl["achieved_A"] / 1e9, "frac": rl["achieved_A"] / HBM_PEAK,
                               "bytes_per_launch": rl["bytes_A"],
                               "traffic": measured_traffic(*args.large, sr)}
        print("large %dx%dx%d: %.2f ms/solve" % (tuple(args.large) + (rl["ms_step"],)), file=sys.stderr,
              flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(*args.size, args.cpu_seconds)

    cfg = summary(args.size, r, args.tol)
    cfg["graphs"] = not args.no_graphs
    cfg["pipelined_polling"] = not args.no_pipeline
    cfg["iteration"] = "single-reduction kernel" if sr else "kernels A + B"
    cfg["large"] = large
    line = {
        "metric": "DOF-matvec/s + % HBM peak, FP32-refined vs FP64 at res 1e-8; p50/p99 ms",
        "value": r["value"], "unit": "DOF-matvec/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 operator / f64 residual",
        "data": "synthetic (seeded densities)",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": r["achieved_A"] / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": r["achieved_A"] / HBM_PEAK,
                     "traffic": measured_traffic(*args.size, sr),
                     "kernel": ("k_hex8_sr<0> (single-reduction PCG iteration: direction, operator, "
                                "residual and solution updates)" if sr else
                                "k_hex8<float,15,7,1> (fused direction update + operator)"),
                     "avg_launch_us": r["avg_A"] * 1e6, "launches": r["launches_A"],
                     "executed_launches": r["executed_A"],
                     "bytes_per_launch": r["bytes_A"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
