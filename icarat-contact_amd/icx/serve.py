"""Serving loop: topology-optimization jobs arriving over time share one GPU.

Jobs arrive as a Poisson stream.  A job is admitted when its device footprint fits
the HBM budget (first come, first served); an admitted job's densities, displacements
and PCG state stay resident until it leaves.  Each design iteration of a job is
filter + projection (icx_opt_step_begin), a mixed-precision solve submitted to the
context's batch (icx_submit), and sensitivity + filter adjoint + OC update
(icx_opt_step_end) once icx_poll reports the solve finished.  The solves of all
admitted jobs advance together (one launch per PCG kernel phase for all of them), and
a job that arrives while others are iterating joins at the next chunk: continuous
batching.  Reported: jobs/s, DOF-matvec/s, and p50/p99 latency of a design iteration
(begin to end, host clock), plus the admission delay.
"""
import collections
import math
import re
import time

import numpy as np

from . import HEXA8, OPT_SINGLE_REDUCTION, PLANE_STRESS, QUAD4, SOLID3D, SOLVE_FUSED, SOLVE_MIXED, IcxError, Job

SEED = 20261205


def job_spec(rng, elems_lo, elems_hi):
    """Element count log-uniform in [elems_lo, elems_hi]; 2D / 3D with probability 1/2;
    2D aspect 3:1 (Q4 plane stress), 3D aspect 2:1:1 (hexa8)."""
    ne = math.exp(rng.uniform(math.log(elems_lo), math.log(elems_hi)))
    if rng.random() < 0.5:
        d0 = max(3, int(round(math.sqrt(3.0 * ne))))
        return (2, (d0, max(1, int(round(d0 / 3.0))), 1))
    d0 = max(2, int(round((4.0 * ne) ** (1.0 / 3.0))))
    d1 = max(1, int(round(d0 / 2.0)))
    return (3, (d0, d1, d1))


def solve_form(ndim):
    """Solve form of a served job: the mixed form for hexa8 jobs, the FP64 fused form for
    quad4 jobs.  A quad4 cantilever of n elements along its length has a Jacobi-scaled
    condition number of order n^2 (6.4e5 at 120 x 40), so FP32 inner iterations lose
    orthogonality: they need about twice the FP64 iteration count at 120 x 40 and
    stall short of 1e-8 from about 840 x 280 up, while the FP64 form converges in
    thousands of iterations; 2D jobs are also too small for the mixed form's halved
    bytes to pay.  Hexa8 jobs of the same element counts are 10 times shorter, and the
    mixed form needs about the FP64 iteration count there."""
    return SOLVE_FUSED if ndim == 2 else SOLVE_MIXED


def footprint_estimate(ndim, div, single_reduction=False, form=None):
    """Device bytes of one served job (cone filter, OC design loop), by its solve form
    (default: solve_form(ndim)).  Mixed form: 84 B per DOF, 92 B per element, 5 B per
    node (constraint mask and the FP32 inverse modulus sums), 432 KB of partial sums
    and state; a hexa8 job on the single-reduction iteration (ICX_OPT_SINGLE_REDUCTION)
    holds one more FP32 vector (88 B per DOF) and 0.375 MB more partial sums.  FP64
    fused form: the same 84 B per DOF (five FP64 vectors in place of the mixed form's
    eight FP32 and one FP64) and 16 B more (node-interleaved copies of x and b),
    88 B per element (no FP32 copy of the moduli), 9 B per node (FP64 inverse modulus
    sums).  icx_job_footprint reports the same buffers once they exist."""
    if form is None:
        form = solve_form(ndim)
    d0, d1, d2 = div
    nn = (d0 + 1) * (d1 + 1) * ((d2 + 1) if ndim == 3 else 1)
    ne = d0 * d1 * (d2 if ndim == 3 else 1)
    n = ndim * nn
    if form == SOLVE_FUSED:
        return 100 * n + 88 * ne + 9 * nn + 432 * 1024
    if ndim == 3 and single_reduction:
        return 88 * n + 92 * ne + 5 * nn + 816 * 1024
    return 84 * n + 92 * ne + 5 * nn + 432 * 1024


def toml_mesh(text):
    """(ndim, div) of a job file's [mesh] section, read only for the footprint estimate
    before admission; the job itself is built by the library's TOML front end."""
    sec = re.search(r"^\s*\[mesh\]\s*$(.*?)(?=^\s*\[|\Z)", text, re.M | re.S)
    if not sec:
        return None
    dim = re.search(r"^\s*dimension\s*=\s*(\d+)", sec.group(1), re.M)
    div = re.search(r"^\s*division\s*=\s*\[([^\]]*)\]", sec.group(1), re.M)
    try:
        d = [int(float(v)) for v in div.group(1).split(",")[:3]]
        return int(dim.group(1)), (d[0], d[1], d[2] if int(dim.group(1)) == 3 else 1)
    except (AttributeError, ValueError, IndexError):
        return None


def create(ctx, spec):
    """The job of a request and its initial design value: spec is (ndim, div) for the
    drawn cantilevers, or the text of a reference job file (TOML) -- mesh, material,
    Dirichlet ranges and tractions read by icx_job_create_toml, x0 = designs_0."""
    if isinstance(spec, str):
        job = Job.from_toml(ctx, spec)
        return job, job.design0
    return make_job(ctx, *spec), None


def make_job(ctx, ndim, div):
    """Cantilever: x = 0 edge/face clamped, unit load at the far end (2D: node
    (L, H/2) in -y; 3D: node (L, W/2, 0) in -z)."""
    d0, d1, d2 = div
    if ndim == 2:
        j = Job(ctx, QUAD4, (d0, d1, 1), (float(d0), float(d1), 1.0), PLANE_STRESS)
        j.add_dirichlet([1, 1, 0, 0, 0, 0], [0] * 6, [0.0, 0.0, 0.0, float(d1), 0.0, 0.0])
        j.make_dof()
        j.add_nodal_load(j.node_index(d0, d1 // 2), 1, -1.0)
    else:
        j = Job(ctx, HEXA8, (d0, d1, d2), (float(d0), float(d1), float(d2)), SOLID3D)
        j.add_dirichlet([1, 1, 1, 0, 0, 0], [0] * 6, [0.0, 0.0, 0.0, float(d1), 0.0, float(d2)])
        j.make_dof()
        j.add_nodal_load(j.node_index(d0, d1 // 2, 0), 2, -1.0)
    return j


def run(ctx, n_jobs=500, rate=20.0, iters=30, elems=(5e3, 1.5e6), budget=240e9, seed=SEED,
        tol=1e-8, maxit=50000, x0=0.4, radius=1.5, specs=None, keep_designs=False, log=None, limit=None,
        cancel=None, max_resident_dofs=None, offset=0, forms="auto", classes=None):
    """Serve n_jobs jobs arriving at `rate` per second (exponential gaps), each running
    `iters` OC design iterations.
    forms: "auto" (solve_form: FP64 fused form for quad4 jobs, mixed form for hexa8
      jobs), or one form (SOLVE_MIXED / SOLVE_FUSED) for every job.
    specs: optional list of requests instead of the drawn sizes -- (ndim, div)
      cantilevers or job-file texts (TOML, the reference's keys; x0 = its designs_0); a
      request the library refuses is recorded under "rejected" and the loop goes on.
    keep_designs: return each job's final design under "designs" and its objective
      history under "objectives".
    limit: serve only the first `limit` jobs of the drawn stream.
    offset: start at job `offset` of the stream (its arrival at t = 0, the later ones at
      their offsets from it): with limit, one segment of a stream too long to serve in
      one sitting, starting empty.
    cancel: {job: design iteration} -- the job is withdrawn while that iteration's solve
      runs (icx_cancel), its budget freed.
    max_resident_dofs: admit a job only while the resident jobs' DOFs of its class, its
      own included, stay within this many (a job larger than the cap is admitted alone).
      The PCG iterations of all resident jobs of a class share each launch, so this
      bounds the work of one batched iteration and with it a design iteration's latency;
      the wait moves to admission (first come, first served within a class).
    classes: DOF bounds splitting the jobs into size classes, e.g. [2e5]: jobs below
      2e5 DOFs and the rest.  Each class has a context of its own (its own stream and
      batch) on ctx's device, served by a host thread of its own, so small jobs' PCG
      iterations no longer share launches with, and wait for, those of large jobs, and
      the classes' kernels overlap on the device.  The HBM budget is shared.  None: one
      class on ctx.
    Returns a dict of the measurements."""
    import threading

    from . import Context

    cancel = dict(cancel or {})
    rng = np.random.default_rng(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate, n_jobs)
    arrival = np.cumsum(gaps) - gaps[0]
    drawn = [job_spec(rng, *elems) for _ in range(n_jobs)]
    specs = list(specs) if specs is not None else drawn
    if offset:
        specs = specs[offset:]
        arrival = arrival[offset:] - arrival[offset]
    if limit is not None:
        specs = specs[:limit]
    n_jobs = len(specs)
    sr = bool(getattr(ctx, "options", {}).get(OPT_SINGLE_REDUCTION, 0))

    def form_of(ndim):
        return solve_form(ndim) if forms == "auto" else forms

    def estimate(spec):
        m = toml_mesh(spec) if isinstance(spec, str) else spec
        return footprint_estimate(*m, single_reduction=sr, form=form_of(m[0])) if m else 0.0

    est = [estimate(s) for s in specs]

    def dofs(spec):
        m = toml_mesh(spec) if isinstance(spec, str) else spec
        if not m:
            return 0
        ndim, (d0, d1, d2) = m
        return ndim * (d0 + 1) * (d1 + 1) * ((d2 + 1) if ndim == 3 else 1)

    ndofs = [dofs(s) for s in specs]
    if max(est) > budget:
        raise ValueError("a job needs %.3g bytes, more than the budget %.3g" % (max(est), budget))
    bounds = sorted(float(b) for b in (classes or []))
    n_cls = len(bounds) + 1
    cls_of = [sum(1 for b in bounds if nd >= b) for nd in ndofs]
    queues = [collections.deque(i for i in range(n_jobs) if cls_of[i] == k) for k in range(n_cls)]
    ctxs = [ctx]
    for _ in range(n_cls - 1):
        c = Context(getattr(ctx, "device", 0))
        for opt, val in getattr(ctx, "options", {}).items():
            c.set_option(opt, val)
        ctxs.append(c)

    lock = threading.Lock()
    sh = {"used": 0.0, "peak_used": 0.0, "peak_active": 0, "active": 0, "peak_resident": 0, "finished": 0,
          "dof_matvecs": 0}
    lat, admit_wait, footprints = [], [], []
    # one record per design iteration: (job, step, DOFs, PCG iterations, FP64
    # replacements, converged, latency ms, resident jobs, resident DOFs, solve form)
    records = []
    designs, histories = {}, {}
    cancelled, rejected = [], []
    errors = []
    # wall time by phase, summed over the classes' threads: batched PCG chunks (poll),
    # per-job work after a solve (step_end: sensitivity, adjoint filter, OC) and before
    # one (step_begin + submit: filter, projection, set-up), admission (create + opt_init)
    splits = [{"poll": 0.0, "step_end": 0.0, "step_begin_submit": 0.0, "admit": 0.0} for _ in range(n_cls)]
    t0 = time.perf_counter()

    def serve_class(k):
        cctx, pending, split = ctxs[k], queues[k], splits[k]
        active = {}
        resident = 0

        def begin(rec):
            tb = time.perf_counter()
            rec["t_begin"] = tb
            rec["job"].opt_step_begin()
            rec["job"].submit(tol=tol, maxit=maxit, form=rec["form"])
            split["step_begin_submit"] += time.perf_counter() - tb

        def leave(rec):
            nonlocal resident
            with lock:
                sh["used"] -= est[rec["spec"]]
                sh["active"] -= 1
                sh["finished"] += 1
            resident -= ndofs[rec["spec"]]

        while pending or active:
            now = time.perf_counter() - t0
            while pending and arrival[pending[0]] <= now:
                i = pending[0]
                if max_resident_dofs is not None and active and resident + ndofs[i] > max_resident_dofs:
                    break
                with lock:  # the budget is reserved before the job is built
                    if sh["used"] + est[i] > budget:
                        break
                    sh["used"] += est[i]
                pending.popleft()
                ta = time.perf_counter()
                try:
                    job, x_init = create(cctx, specs[i])
                except IcxError as e:
                    rejected.append((i, str(e)))
                    with lock:
                        sh["used"] -= est[i]
                        sh["finished"] += 1
                    continue
                with lock:
                    sh["active"] += 1
                    sh["peak_used"] = max(sh["peak_used"], sh["used"])
                    sh["peak_active"] = max(sh["peak_active"], sh["active"])
                form = form_of(job.ndim)
                job.opt_init(np.full(job.nelx, x0 if x_init is None else x_init), driver=0, radius=radius,
                             projection=True, form=form, tol=tol, maxit=maxit)
                split["admit"] += time.perf_counter() - ta
                rec = {"job": job, "spec": i, "step": 0, "numeq": job.numeq, "form": form}
                active[id(job)] = rec
                resident += ndofs[i]
                with lock:
                    sh["peak_resident"] = max(sh["peak_resident"], resident)
                admit_wait.append(time.perf_counter() - t0 - arrival[i])
                begin(rec)
            if active:
                tp = time.perf_counter()
                finished = cctx.poll(chunks=1)
                split["poll"] += time.perf_counter() - tp
                for rec in [r for r in active.values() if cancel.get(r["spec"]) == r["step"]]:
                    if rec["job"] in finished:
                        continue
                    rec["job"].cancel()
                    rec["job"].close()
                    del active[id(rec["job"])]
                    leave(rec)
                    cancelled.append(rec["spec"])
                for job in finished:
                    rec = active[id(job)]
                    info = job.last_info()
                    te = time.perf_counter()
                    step_rec = job.opt_step_end(info)
                    split["step_end"] += time.perf_counter() - te
                    if keep_designs:
                        histories.setdefault(rec["spec"], []).append(step_rec["objective"])
                    lat.append(time.perf_counter() - rec["t_begin"])
                    records.append((rec["spec"], rec["step"], rec["numeq"], info["iterations"], info["refinements"],
                                    bool(info["converged"]), lat[-1] * 1e3, len(active), resident, rec["form"]))
                    with lock:
                        sh["dof_matvecs"] += rec["numeq"] * info["matvecs"]
                    rec["step"] += 1
                    if rec["step"] < iters:
                        begin(rec)
                    else:
                        if len(footprints) < 8:
                            footprints.append((job.footprint(), est[rec["spec"]]))
                        if keep_designs:
                            designs[rec["spec"]] = job.opt_design()
                        job.close()
                        del active[id(job)]
                        leave(rec)
                        if log:
                            log("job %d done (%d left)" % (rec["spec"], n_jobs - sh["finished"]))
            elif pending:  # next arrival, or budget held by the other classes
                time.sleep(max(0.001, min(0.01, arrival[pending[0]] - (time.perf_counter() - t0))))

    def guarded(k):
        try:
            serve_class(k)
        except BaseException as e:  # re-raised by the caller once every class has stopped
            errors.append(e)
            for q in queues:  # the other classes admit nothing more
                q.clear()

    if n_cls == 1:
        serve_class(0)
    else:
        threads = [threading.Thread(target=guarded, args=(k,), daemon=True) for k in range(n_cls)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for c in ctxs[1:]:
            c.synchronize()
            c.close()
        if errors:
            raise errors[0]
    wall = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    return {
        "jobs": n_jobs, "design_iterations": len(lat), "wall_s": wall,
        "jobs_per_s": n_jobs / wall, "dof_matvec_per_s": sh["dof_matvecs"] / wall,
        "iteration_ms_p50": float(np.percentile(lat_ms, 50)),
        "iteration_ms_p99": float(np.percentile(lat_ms, 99)),
        "admission_wait_ms_p50": float(np.percentile(admit_wait, 50) * 1e3),
        "admission_wait_ms_p99": float(np.percentile(admit_wait, 99) * 1e3),
        "peak_admitted_jobs": sh["peak_active"], "peak_budget_bytes": sh["peak_used"],
        "arrival_rate_per_s": rate, "iterations_per_job": iters, "elements": list(elems),
        "stream_jobs_drawn": len(drawn),
        "budget_bytes": budget, "tol": tol,
        "footprint_checks": footprints,
        "cancelled": cancelled,
        "rejected": rejected,
        "wall_split_s": {key: sum(sp[key] for sp in splits) for key in splits[0]},
        "offset": offset,
        "size_classes_dofs": bounds,
        "latencies_ms": lat_ms.tolist(),
        "iteration_records": records,
        "unconverged_solves": sum(1 for r in records if not r[5]),
        "admission_wait_ms": (np.array(admit_wait) * 1e3).tolist(),
        "designs": designs if keep_designs else None,
        "objectives": histories if keep_designs else None,
        "max_used_within_budget": sh["peak_used"] <= budget,
        "max_resident_dofs": max_resident_dofs, "peak_resident_dofs": sh["peak_resident"],
    }
