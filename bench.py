"""Benchmark of the FE assembly + Jacobi-PCG hot path on one MI355X.

Workload: 64^3 Kuhn cube (1,572,864 tets), interior nodes jittered +-0.2h,
P2 variable-coefficient diffusion with per-element lognormal k, 11-point
quadrature, nodal f ~ U[-1,1], homogeneous Dirichlet on all faces.

A step is one numeric assembly of the full FP64 matrix (the values
traitsMatrixBuilder::buildEpetraCrs produces) into the fixed CSR pattern, with
mesh, coefficients, pattern and incidence records already resident in HBM.
`value` = elements assembled per second (M elem/s) over the K timed steps.
The same JSON line also reports the PCG times to 1e-10 (FP64 plain and FP32
inner + FP64 refinement), the gradient/forward cost ratio, the roofline of the
assembly and of one PCG iteration, and the CPU baseline (the FP64 C oracle on
a bounded sample of the same mesh, one host core).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from morganapublic_amd import _lib as L  # noqa: E402
from morganapublic_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0


class HipTimer:
    """hipEvent pair on a given stream (the library's stream)."""

    def __init__(self, stream: int):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.stream = C.c_void_p(stream)
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.e0)) == 0
        assert self.hip.hipEventCreate(C.byref(self.e1)) == 0

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def to_device(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p.value

    def zeros(self, nbytes: int) -> int:
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) == 0
        assert self.hip.hipMemset(p, 0, C.c_size_t(nbytes)) == 0
        return p.value

    def free(self, ptr: int):
        self.hip.hipFree.argtypes = [C.c_void_p]
        assert self.hip.hipFree(C.c_void_p(ptr)) == 0

    def to_host(self, ptr: int, n: int, dtype=np.float64) -> np.ndarray:
        a = np.zeros(n, dtype=dtype)
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                  C.c_size_t(a.nbytes), 2) == 0
        return a

    def start(self):
        assert self.hip.hipEventRecord(self.e0, self.stream) == 0

    def stop_ms(self) -> float:
        assert self.hip.hipEventRecord(self.e1, self.stream) == 0
        assert self.hip.hipEventSynchronize(self.e1) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1) == 0
        return float(ms.value)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def bytes_assembly(T, V, nloc, nnz_full, vb):
    # compulsory traffic: element dofs (int32) + one coefficient per tet +
    # vertex coordinates once + every stored value once
    return T * (4 * nloc + 8) + 24 * V + vb * nnz_full


def bytes_pcg_iter(n, nnz, vb):
    return nnz * (vb + 4) + 4 * (n + 1) + 13 * vb * n


def bytes_adjoint(T, V, nloc, n_sys, n_full):
    """Algorithmic bytes of the adjoint element kernel: per element its 4
    vertex ids and nloc dof ids, one gradient written; coordinates once; u and
    lambda once each (system numbering) and the dof -> system map once."""
    return T * (16 + 4 * nloc + 8) + 24 * V + 16 * n_sys + 4 * n_full


def bytes_spmv(n, nnz, vb):
    """Algorithmic bytes of y = A x: values and 4-byte columns once, row
    pointers, x and y once."""
    return nnz * (vb + 4) + 4 * (n + 1) + 2 * vb * n


TRAFFIC_PASSES = (
    ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"),
    ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum"),
)
TRAFFIC_REPS = 5


def traffic_child(n):
    """Run under rocprofv3 --pmc by measure_traffic: the FP64 assembly of the
    full n^3 P2 system, TRAFFIC_REPS times after a short settle."""
    p = W.config_problem(2, n=n)
    m = p.mesh
    ctx = L.Context(0)
    ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
    ctx.dofmap_build(2)
    ctx.pattern_build(L.SYS_FULL)
    ctx.field_set(L.FIELD_STIFF, p.k)
    for _ in range(TRAFFIC_REPS):
        ctx.assemble(L.SYS_FULL, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=11)
    ctx.synchronize()
    ctx.close()


def measure_traffic(n, timeout=150):
    """Bytes between L2 and memory of one FP64 assembly, measured in this run:
    two rocprofv3 counter passes over a child process (traffic_child), one per
    counter group (request counts by size: read = 128 RDREQ_128B + 64
    RDREQ_64B + 32 RDREQ_32B, write = 64 WRREQ_64B + 32 (WRREQ - WRREQ_64B)),
    summed over the assembly kernels (median per kernel over the dispatches).
    None (with the reason) when the profiler is unavailable or fails."""
    import csv
    import shutil
    import statistics
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    env = dict(os.environ, TMPDIR="/tmp")
    per = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for k, counters in enumerate(TRAFFIC_PASSES):
            out = os.path.join(td, "p%d" % k)
            cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", "pmc",
                   "--", sys.executable, os.path.abspath(__file__), "--traffic-child",
                   "--n", str(n)]
            # own process group, so that a pass past its time limit is ended
            # together with the program it profiles
            pr = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                                  stderr=subprocess.DEVNULL, start_new_session=True)
            try:
                rc = pr.wait(timeout=timeout)
            except subprocess.TimeoutExpired:
                os.killpg(pr.pid, 9)
                pr.wait()
                return None, "rocprofv3 pass %d timed out" % k
            files = []
            for root, _, names in os.walk(out):
                files += [os.path.join(root, f) for f in names if f.endswith("counter_collection.csv")]
            if not files:
                return None, "rocprofv3 pass %d: no counter file (rc %d)" % (k, rc)
            for f in files:
                with open(f) as fh:
                    for row in csv.DictReader(fh):
                        name = row["Kernel_Name"]
                        if "k_asm_rows" not in name and "k_elem_factors" not in name:
                            continue
                        key = name[:name.find("(")] if "(" in name else name
                        per.setdefault(key, {}).setdefault(row["Counter_Name"], []).append(
                            float(row["Counter_Value"]))
    if not per:
        return None, "no assembly kernel in the counter files"
    kernels, rd_tot, wr_tot = [], 0.0, 0.0
    for key, cs in per.items():
        med = {c: statistics.median(v) for c, v in cs.items()}
        calls = max(len(v) for v in cs.values())
        rd = (128 * med.get("TCC_EA0_RDREQ_128B_sum", 0) + 64 * med.get("TCC_EA0_RDREQ_64B_sum", 0)
              + 32 * med.get("TCC_EA0_RDREQ_32B_sum", 0))
        wr64 = med.get("TCC_EA0_WRREQ_64B_sum", 0)
        wr = 64 * wr64 + 32 * (med.get("TCC_EA0_WRREQ_sum", 0) - wr64)
        # launches per assembly: dispatches of this kernel over the repetitions
        per_asm = calls / TRAFFIC_REPS
        rd_tot += rd * per_asm
        wr_tot += wr * per_asm
        kernels.append({"kernel": key[:160], "launches_per_assembly": per_asm,
                        "read_bytes": int(rd), "write_bytes": int(wr)})
    return {"bytes_per_assembly": int(rd_tot + wr_tot), "read_bytes": int(rd_tot),
            "write_bytes": int(wr_tot), "kernels": kernels}, None


def bench_config5(ctx, timer, keep, n=128, solve=True):
    """128^3 P1, lognormal k: the assembly alone (FP64 and FP32, interior
    system), then (solve) forward + adjoint solve with both FP32 + FP64
    refinement methods to 1e-10, vs the FP64 plain method; dJ/dk for
    J = ||u - u_obs||^2."""
    p = W.config_problem(5, n=n)
    m = p.mesh
    ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
    ctx.dofmap_build(1)
    ctx.dirichlet_set(list(p.dirichlet_geo))
    nint, nnz_int = ctx.pattern_build(L.SYS_INTERIOR)
    ctx.field_set(L.FIELD_STIFF, p.k)
    fixed = ctx.dirichlet_mask()
    db = timer.to_device(np.ascontiguousarray(p.f[fixed == 0]))
    uobs = (0.01 * (2 * W.stream_uniform(0, W.STREAM_UOBS, fixed.size) - 1))[fixed == 0]
    du, dl, dg = timer.zeros(8 * nint), timer.zeros(8 * nint), timer.zeros(8 * m.nT)
    out = {"tets": m.nT, "dofs_interior": nint, "nnz_interior": nnz_int}
    # the P1 assembly alone (interior system; algorithmic bytes with nnz_int)
    kw = dict(stiff_coeff=L.COEFF_P0, stiff_rule=1)
    b_asm = m.nT * (4 * 4 + 8) + 24 * m.nV + 8 * nnz_int
    for prec, name in ((L.PREC_F64, "f64"), (L.PREC_F32, "f32")):
        # the same device settle as the headline step (about 0.3 s of the
        # step itself, untimed): measured right after the symbolic phase,
        # the first assemblies run up to 25% slower
        t_settle = time.perf_counter()
        while time.perf_counter() - t_settle < 0.3:
            for _ in range(10):
                ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, prec=prec, **kw)
            timer.sync()
        timer.start()
        for _ in range(20):
            ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, prec=prec, **kw)
        ms = timer.stop_ms() / 20
        out["assembly_%s_us" % name] = round(1e3 * ms, 1)
        if prec == L.PREC_F64:
            out["assembly_f64_roofline_frac"] = round(b_asm / (ms * 1e-3) / 8e12, 4)
            rp5, ci5 = ctx.pattern(L.SYS_INTERIOR)
            keep["p1_128"] = (p, rp5, ci5, ctx.values(L.SYS_INTERIOR), ctx.dirichlet_mask())
        log("cfg5 P1 assembly %s: %.1f us (%.0f M elem/s)" % (name, 1e3 * ms, m.nT / ms / 1e3))
    if not solve:
        for ptr in (db, du, dl, dg):
            timer.free(ptr)
        return out
    grads = {}
    for name, mode in (("f64", L.SOLVE_F64), ("f32_refine", L.SOLVE_F32_REFINE),
                       ("f32_reliable", L.SOLVE_F32_RELIABLE)):
        kw = dict(stiff_coeff=L.COEFF_P0, stiff_rule=1)
        timer.sync()
        timer.start()
        ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, **kw)
        if mode != L.SOLVE_F64:
            ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, prec=L.PREC_F32, **kw)
        fi = ctx.pcg_device(db, du, mode=mode, tol=1e-10)
        ms_fwd = timer.stop_ms()
        drhs = timer.to_device(2.0 * (timer.to_host(du, nint) - uobs))
        timer.sync()
        timer.start()
        ai = ctx.pcg_device(drhs, dl, mode=mode, tol=1e-10)
        ctx.adjoint_grad_k_device(L.SYS_INTERIOR, du, dl, dg, rule=1)
        ms_adj = timer.stop_ms()
        grads[name] = timer.to_host(dg, m.nT)
        timer.free(drhs)
        out[name] = {"forward_ms": round(ms_fwd, 2), "gradient_ms": round(ms_adj, 2),
                     "grad_fwd_ratio": round(ms_adj / ms_fwd, 3),
                     "forward_iters": fi["iters"], "adjoint_iters": ai["iters"],
                     "forward_relres": fi["relres"], "adjoint_relres": ai["relres"]}
        log("cfg5 %s: forward %.1f ms (%d it), gradient %.1f ms (%d it)"
            % (name, ms_fwd, fi["iters"], ms_adj, ai["iters"]))
    b = grads["f64"]
    for name in ("f32_refine", "f32_reliable"):
        out["grad_rel_err_%s_vs_f64" % name] = float(np.linalg.norm(grads[name] - b) / np.linalg.norm(b))
    for ptr in (db, du, dl, dg):
        timer.free(ptr)
    return out


def bytes_dd(nT, nV, vjp):
    """Algorithmic bytes of the drift-diffusion residual (+ VJP): per element
    the 4 vertex ids, eps, rho, mu and the 6 per-element dual areas once; per
    vertex its coordinates, phi, n and the fixed flag once, and the two
    residuals written once; the VJP also writes d/d(eps, mu) per element (its
    residuals are recomputed, not re-read)."""
    b = nT * (16 + 24 + 48) + nV * (24 + 16 + 1 + 16)
    return b + (nT * 16 if vjp else 0)


def bench_config4(ctx, timer, n=96, reps=5):
    """96^3 P1 drift-diffusion: dual-mesh data, SG + Poisson residual with its
    loss, and the FP64 VJP w.r.t. (eps, mu), device buffers."""
    p = W.config_drift_diffusion(n)
    m = p.mesh
    ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
    ctx.dofmap_build(1)
    ctx.dirichlet_set(list(p.dirichlet_geo))
    fixed = ctx.dirichlet_mask()
    p.fix_state(fixed)
    timer.sync()
    t0 = time.perf_counter()
    ctx.dual_fv_build()
    timer.sync()
    t_dual = time.perf_counter() - t0
    dev = [timer.to_device(a) for a in (p.eps, p.rho, p.mu, p.phi, p.dens)]
    dfix = timer.to_device(fixed.astype(np.uint8))
    rp, rn, loss = timer.zeros(8 * m.nV), timer.zeros(8 * m.nV), timer.zeros(8)
    ge, gm = timer.zeros(8 * m.nT), timer.zeros(8 * m.nT)
    args = dev + [dfix]
    ctx.dd_residual_device(*args, rp, rn, loss)
    timer.sync()
    timer.start()
    for _ in range(reps):
        ctx.dd_residual_device(*args, rp, rn, loss)
    ms_res = timer.stop_ms() / reps
    timer.start()
    for _ in range(reps):
        ctx.dd_vjp_device(*args, rp, rn, ge, gm)
    ms_vjp = timer.stop_ms() / reps
    log("cfg4: dual data %.1f ms, residual+loss %.3f ms, residual+VJP %.3f ms"
        % (t_dual * 1e3, ms_res, ms_vjp))
    for ptr in args + [rp, rn, loss, ge, gm]:
        timer.free(ptr)
    b_res, b_vjp = bytes_dd(m.nT, m.nV, False), bytes_dd(m.nT, m.nV, True)
    return {"tets": m.nT, "vertices": m.nV, "dual_build_ms": round(t_dual * 1e3, 2),
            "residual_loss_ms": round(ms_res, 3), "residual_vjp_ms": round(ms_vjp, 3),
            "residual_loss_bytes": b_res, "residual_vjp_bytes": b_vjp,
            "residual_loss_roofline_frac": round(b_res / (ms_res * 1e-3) / 8e12, 4),
            "residual_vjp_roofline_frac": round(b_vjp / (ms_vjp * 1e-3) / 8e12, 4)}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def oracle_parity(keep, threads):
    """Max row-scaled difference between the GPU's FP64 values and the
    oracle's element matrices added into the same pattern on host threads
    (O.accumulate), at the full 64^3 P2 and 128^3 P1 sizes; the oracle is
    the checker here, the values checked are the ones the timed steps made."""
    import oracle as O
    out = {}

    def rowscaled(rp, got, want):
        rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
        sc = np.zeros(rp.size - 1)
        np.maximum.at(sc, rows, np.abs(want))
        sc[sc == 0] = np.abs(want).max()
        return float((np.abs(got - want) / sc[rows]).max())

    if "p2_64" in keep:
        p, rp, ci, got = keep["p2_64"]
        d = O.elem_dofs(2, p.mesh)
        want, miss = O.accumulate(2, p.mesh, O.OP_STIFF, d, rp, ci, c_stiff=p.k, rule_stiff=4,
                                  threads=threads)
        out["p2_64_full_rowscaled_max"] = rowscaled(rp, got, want) if miss == 0 else None
        out["p2_64_full_missing_pairs"] = int(miss)
    if "p1_128" in keep:
        p, rp, ci, got, fixed = keep["p1_128"]
        sysmap = np.where(fixed == 0, np.cumsum(fixed == 0) - 1, -1).astype(np.int32)
        d = sysmap[O.elem_dofs(1, p.mesh)]
        want, miss = O.accumulate(1, p.mesh, O.OP_STIFF, d, rp, ci, c_stiff=p.k, rule_stiff=1,
                                  threads=threads)
        out["p1_128_interior_rowscaled_max"] = rowscaled(rp, got, want) if miss == 0 else None
        out["p1_128_interior_missing_pairs"] = int(miss)
    out["threads"] = threads
    log("oracle parity at full size: %s" % out)
    return out


def cpu_pcg_baseline(keep, threads, gpu_iters):
    """The oracle's Jacobi-PCG iteration (or_pcg_fixed) on the 64^3 P2
    interior system, SpMV and vector updates split over host threads; timed
    per iteration (a fixed count from x = 0), on all threads and on one, and
    scaled to the FP64 GPU solve's iteration count to 1e-10."""
    import oracle as O
    if "p2_64_int" not in keep:
        return None
    rp, ci, va, b = keep["p2_64_int"]
    _, t_n, _ = O.pcg_fixed(rp, ci, va, b, 20, threads=threads)
    _, t_1, _ = O.pcg_fixed(rp, ci, va, b, 3, threads=1)
    out = {"ms_per_iter": round(t_n * 1e3, 3), "threads": threads,
           "ms_per_iter_1core": round(t_1 * 1e3, 3),
           "iters_timed": 20, "gpu_f64_iters": gpu_iters,
           "est_ms_to_1e-10": round(t_n * 1e3 * gpu_iters, 1) if gpu_iters else None,
           "kind": "port",
           "sample": "oracle Jacobi-PCG iteration (or_pcg_fixed) on the 64^3 P2 interior system "
                     "(N = %d, nnz = %d): 20 iterations on %d threads, 3 on one core; the "
                     "time to 1e-10 is the per-iteration time x the GPU FP64 solve's iteration "
                     "count" % (b.size, va.size, threads)}
    log("CPU PCG baseline: %.2f ms/iteration on %d threads (%.2f on one core)"
        % (t_n * 1e3, threads, t_1 * 1e3))
    return out


def cpu_baseline(problem, budget_s=20.0, max_threads=16):
    """FP64 oracle (the reference algorithm restated in C: per-quadrature-point
    geometry, COO insert + sort/merge) on the first elements of the same mesh.
    One core first (the sample grows until it takes a few seconds), then one
    thread per host core (at most max_threads, the host share of one GPU),
    each assembling its own contiguous block of elements into its own CSR, as
    the ranks of the reference's MPI assembly do (no cross-thread merge); the
    oracle's calls release the GIL.  About budget_s of CPU work in all."""
    import concurrent.futures as cf
    import oracle as O
    m = problem.mesh
    O.lib()

    def run(t0_, t1_):
        sub = W.TetMesh(n=m.n, xyz=m.xyz, tets=m.tets[t0_:t1_], btri=m.btri[:0], bgeo=m.bgeo[:0])
        O.assemble(2, sub, O.OP_STIFF, c_stiff=problem.k[t0_:t1_], rule_stiff=4)

    ntake = 2000
    while True:
        t0 = time.perf_counter()
        run(0, ntake)
        dt1 = time.perf_counter() - t0
        if dt1 > budget_s / 4 or ntake >= m.nT:
            break
        ntake = min(m.nT, int(ntake * max(2.0, (budget_s / 4) / max(dt1, 1e-3))))
    rate1 = ntake / dt1
    nthreads = max(1, min(max_threads, os.cpu_count() or 1))
    per = int(min(m.nT // nthreads, max(1000, rate1 * (0.75 * budget_s) / nthreads)))
    with cf.ThreadPoolExecutor(nthreads) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda t: run(t * per, (t + 1) * per), range(nthreads)))
        dtn = time.perf_counter() - t0
    rate = nthreads * per / dtn
    return {"value": rate / 1e6, "unit": "M elem/s", "cores": nthreads, "kind": "port",
            "sample": "FP64 C oracle (per-point geometry, COO sort-merge, 4-point rule) on the "
                      "64^3 P2 mesh: %d threads x %d tets in %.2f s; 1 core: first %d tets in "
                      "%.2f s (%.4f M elem/s); host %s, nproc %d"
                      % (nthreads, per, dtn, ntake, dt1, rate1 / 1e6, _cpu_model(),
                         os.cpu_count() or 0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes that measure the assembly's memory traffic")
    ap.add_argument("--assembly-only", action="store_true",
                    help="the 64^3 P2 and 128^3 P1 assemblies only (no solves, extras or CPU baseline)")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 96^3 drift-diffusion and 128^3 gradient extras")
    ap.add_argument("--dd-only", action="store_true",
                    help="the 96^3 drift-diffusion residual and VJP only (--steps repetitions "
                         "each), printed as one JSON line")
    args = ap.parse_args()
    if args.traffic_child:
        traffic_child(args.n)
        return
    if args.dd_only:
        ctx = L.Context(0)
        print(json.dumps(bench_config4(ctx, HipTimer(ctx.stream), n=96, reps=args.steps)))
        return
    if args.assembly_only:
        args.no_solve = args.no_cpu = True
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    t_setup = time.perf_counter()
    p = W.config_problem(2, n=args.n)
    m = p.mesh
    log("mesh n=%d T=%d V=%d generated in %.1fs" % (args.n, m.nT, m.nV, time.perf_counter() - t_setup))
    ctx = L.Context(0)
    timer = HipTimer(ctx.stream)
    ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
    t0 = time.perf_counter()
    ndofs = ctx.dofmap_build(2)
    nrows, nnz = ctx.pattern_build(L.SYS_FULL)
    nfree = ctx.dirichlet_set(list(p.dirichlet_geo))
    nint, nnz_int = ctx.pattern_build(L.SYS_INTERIOR)
    timer.sync()
    t_sym = time.perf_counter() - t0
    log("symbolic (edges, dofs, 2 patterns, slot maps, incidence) %.3fs: N=%d nnz=%d N_int=%d "
        "nnz_int=%d" % (t_sym, ndofs, nnz, nint, nnz_int))
    ctx.field_set(L.FIELD_STIFF, p.k)
    keep = {}

    def assemble(sysid=L.SYS_FULL, prec=L.PREC_F64):
        ctx.assemble(sysid, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=11, prec=prec)

    # device settle before the warm-up steps: the same step, untimed, for
    # about 0.3 s (the first ~60 steps after the symbolic phase run up to 10%
    # slower while the device clocks come up; a few warm-up steps do not
    # cover that)
    settle, t_settle = 0, time.perf_counter()
    while time.perf_counter() - t_settle < 0.3 or settle < 50:
        for _ in range(10):
            assemble()
        settle += 10
        timer.sync()
    for _ in range(args.warmup):
        assemble()
    timer.sync()
    w0 = time.perf_counter()
    timer.start()
    for _ in range(args.steps):
        assemble()
    ms_total = timer.stop_ms()
    timer.sync()
    wall = time.perf_counter() - w0
    ms_step = ms_total / args.steps
    value = m.nT / (ms_step * 1e-3) / 1e6
    log("assembly FP64: %.3f ms/step (wall %.3f ms/step), %.1f M elem/s"
        % (ms_step, wall * 1e3 / args.steps, value))
    b_asm = bytes_assembly(m.nT, m.nV, 10, nnz, 8)
    achieved = b_asm / (ms_step * 1e-3) / 1e9
    if not args.no_cpu:
        rp_f, ci_f = ctx.pattern(L.SYS_FULL)
        keep["p2_64"] = (p, rp_f, ci_f, ctx.values(L.SYS_FULL))

    # attainable HBM bandwidth: a plain 1 GiB device copy on the same stream
    cbytes = 1 << 30
    src, dst = timer.zeros(cbytes), timer.zeros(cbytes)
    ctx.copy_device(dst, src, cbytes)
    timer.sync()
    timer.start()
    for _ in range(5):
        ctx.copy_device(dst, src, cbytes)
    ms_copy = timer.stop_ms() / 5
    copy_gbs = 2 * cbytes / (ms_copy * 1e-3) / 1e9
    log("device copy 1 GiB: %.3f ms, %.0f GB/s (read + write)" % (ms_copy, copy_gbs))
    timer.free(src)
    timer.free(dst)

    # FP32 assembly (inner-solve matrix)
    assemble(prec=L.PREC_F32)
    timer.sync()
    timer.start()
    for _ in range(args.steps):
        assemble(prec=L.PREC_F32)
    ms_asm32 = timer.stop_ms() / args.steps

    out_extra = {}
    if not args.no_solve:
        assemble(L.SYS_INTERIOR)
        assemble(L.SYS_INTERIOR, L.PREC_F32)
        fixed = ctx.dirichlet_mask()
        b = np.ascontiguousarray(p.f[fixed == 0])
        db = timer.to_device(b)
        dx = timer.zeros(8 * nint)
        # SpMV alone on the interior system (the PCG's operator), both precisions
        spmv = {}
        xs = {L.PREC_F64: db, L.PREC_F32: timer.zeros(4 * nint)}
        ys = {L.PREC_F64: dx, L.PREC_F32: timer.zeros(4 * nint)}
        for name, prec, vb in (("f64", L.PREC_F64, 8), ("f32", L.PREC_F32, 4)):
            ctx.spmv_device(L.SYS_INTERIOR, xs[prec], ys[prec], prec=prec)
            timer.sync()
            timer.start()
            for _ in range(20):
                ctx.spmv_device(L.SYS_INTERIOR, xs[prec], ys[prec], prec=prec)
            us = timer.stop_ms() / 20 * 1e3
            gbs = bytes_spmv(nint, nnz_int, vb) / (us * 1e-6) / 1e9
            spmv[name] = {"us": round(us, 1), "GBps": round(gbs, 1),
                          "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
            log("SpMV %s (interior): %.1f us, %.0f GB/s on algorithmic bytes" % (name, us, gbs))
        timer.free(xs[L.PREC_F32])
        timer.free(ys[L.PREC_F32])
        res = {}
        modes = (("f64", L.SOLVE_F64), ("f32_refine", L.SOLVE_F32_REFINE),
                 ("f32_reliable", L.SOLVE_F32_RELIABLE))
        for name, mode in modes:
            ctx.pcg_device(db, dx, mode=mode, tol=1e-10)  # warm-up
            best = None
            for _ in range(2):
                timer.sync()
                timer.start()
                info = ctx.pcg_device(db, dx, mode=mode, tol=1e-10)
                ms = timer.stop_ms()
                if best is None or ms < best[0]:
                    best = (ms, info)
            res[name] = {"ms": round(best[0], 3), "iters": best[1]["iters"],
                         "refinements": best[1]["outer"], "relres": best[1]["relres"]}
            log("PCG %s: %.2f ms, %d iters, %d refinements, relres %.2e"
                % (name, best[0], best[1]["iters"], best[1]["outer"], best[1]["relres"]))
        # the faster FP32 + FP64-refinement method is the mixed solver of record
        mixed = min(("f32_refine", "f32_reliable"), key=lambda k: res[k]["ms"])
        mmode = dict(modes)[mixed]
        # config 3 errors (outside the timed regions): FP32 matrix against the
        # FP64 one (the parity tests pin FP64 to the oracle), and the mixed
        # solution against the FP64 solution
        rp_i, _ = ctx.pattern(L.SYS_INTERIOR)
        v64, v32 = ctx.values(L.SYS_INTERIOR), ctx.values(L.SYS_INTERIOR, prec=L.PREC_F32)
        rows = np.repeat(np.arange(nint), np.diff(rp_i))
        rmax = np.zeros(nint)
        np.maximum.at(rmax, rows, np.abs(v64))
        mat_err = float((np.abs(v32.astype(np.float64) - v64) / rmax[rows]).max())
        if not args.no_cpu:
            keep["p2_64_int"] = (rp_i, ctx.pattern(L.SYS_INTERIOR)[1], v64, b)
        del v32, rows
        ctx.pcg_device(db, dx, mode=L.SOLVE_F64, tol=1e-10)
        x64 = timer.to_host(dx, nint)
        ctx.pcg_device(db, dx, mode=mmode, tol=1e-10)
        x32 = timer.to_host(dx, nint)
        sol_err = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
        log("config 3: FP32 matrix max row-scaled error %.2e, mixed solution vs FP64 %.2e"
            % (mat_err, sol_err))
        # gradient / forward cost: forward = assemble + mixed solve; gradient =
        # adjoint mixed solve + adjoint element kernel (dJ/dk, J = ||u - u_obs||^2)
        uobs = (2 * W.stream_uniform(0, W.STREAM_UOBS, fixed.size) - 1)[fixed == 0]
        du = timer.zeros(8 * nint)
        dlam = timer.zeros(8 * nint)
        dg = timer.zeros(8 * m.nT)
        timer.sync()
        timer.start()
        assemble(L.SYS_INTERIOR)
        assemble(L.SYS_INTERIOR, L.PREC_F32)
        ctx.pcg_device(db, du, mode=mmode, tol=1e-10)
        ms_fwd = timer.stop_ms()
        drhs = timer.to_device(2.0 * (timer.to_host(du, nint) - uobs))
        timer.sync()
        timer.start()
        ctx.pcg_device(drhs, dlam, mode=mmode, tol=1e-10)
        ctx.adjoint_grad_k_device(L.SYS_INTERIOR, du, dlam, dg, rule=11)
        ms_grad = timer.stop_ms()
        log("forward %.2f ms, gradient %.2f ms, ratio %.3f" % (ms_fwd, ms_grad, ms_grad / ms_fwd))
        # the adjoint element kernel alone (dJ/dk_e = -lam_e^T Khat_e u_e).
        # One call is shorter than the host's time to issue it, so back-to-back
        # calls would time the host: each timed call is queued behind a 1 GiB
        # device copy (the stream is busy while the events and the kernel are
        # issued), and the median of 7 such single-call times is taken
        cb = 1 << 30
        csrc, cdst = timer.zeros(cb), timer.zeros(cb)
        t_adj = []
        for _ in range(7):
            ctx.copy_device(cdst, csrc, cb)
            timer.start()
            ctx.adjoint_grad_k_device(L.SYS_INTERIOR, du, dlam, dg, rule=11)
            t_adj.append(timer.stop_ms())
        timer.free(csrc)
        timer.free(cdst)
        us_adj = float(np.median(t_adj)) * 1e3
        b_adj = bytes_adjoint(m.nT, m.nV, 10, nint, ndofs)
        log("adjoint element kernel: %.1f us, %.0f GB/s on %d algorithmic bytes"
            % (us_adj, b_adj / (us_adj * 1e-6) / 1e9, b_adj))
        it_bytes64 = bytes_pcg_iter(nint, nnz_int, 8)
        it_bytes32 = bytes_pcg_iter(nint, nnz_int, 4)
        out_extra = {
            "spmv_interior": spmv,
            "config3_errors": {"fp32_matrix_rowscaled_max": mat_err,
                               "mixed_solution_rel_vs_fp64": sol_err,
                               "fp32_element_matrices": "closed form in FP32 VALU (no MFMA: "
                                                        "no dense product is left once the "
                                                        "quadrature is contracted)"},
            "pcg_to_1e-10": res,
            "pcg_mixed_method": mixed,
            "pcg_speedup_mixed_vs_f64": round(res["f64"]["ms"] / res[mixed]["ms"], 3),
            "pcg_speedup_f32_refine_vs_f64": round(res["f64"]["ms"] / res["f32_refine"]["ms"], 3),
            "pcg_speedup_f32_reliable_vs_f64": round(res["f64"]["ms"] / res["f32_reliable"]["ms"], 3),
            "grad_fwd_ratio": round(ms_grad / ms_fwd, 3),
            "adjoint_kernel": {"us": round(us_adj, 1), "bytes": b_adj,
                               "GBps": round(b_adj / (us_adj * 1e-6) / 1e9, 1),
                               "roofline_frac": round(b_adj / (us_adj * 1e-6) / 8e12, 4)},
            "forward_ms": round(ms_fwd, 3), "gradient_ms": round(ms_grad, 3),
            "pcg_f64_GBps_per_iter": round(it_bytes64 * res["f64"]["iters"] / (res["f64"]["ms"] * 1e-3) / 1e9, 1),
            "pcg_f64_roofline_frac": round(it_bytes64 * res["f64"]["iters"] / (res["f64"]["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        }

    if not args.no_solve and not args.no_configs:
        out_extra["config5_gradient_128"] = bench_config5(ctx, timer, keep if not args.no_cpu else {})
        out_extra["config4_drift_diffusion_96"] = bench_config4(ctx, timer)
    elif args.assembly_only and not args.no_configs:
        out_extra["config5_gradient_128"] = bench_config5(ctx, timer, {}, solve=False)
    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not args.no_traffic:
        traffic, traffic_note = measure_traffic(args.n)
        if traffic:
            log("assembly traffic (L2 <-> memory, counted in this run): %.1f MB read + %.1f MB "
                "written per FP64 assembly" % (traffic["read_bytes"] / 1e6, traffic["write_bytes"] / 1e6))
        else:
            log("assembly traffic not measured: %s" % traffic_note)
    cpu = None
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        cpu = cpu_baseline(p)
        out_extra["oracle_parity"] = oracle_parity(keep, threads)
        gpu_iters = out_extra.get("pcg_to_1e-10", {}).get("f64", {}).get("iters")
        out_extra["cpu_pcg_baseline"] = cpu_pcg_baseline(keep, threads, gpu_iters)
    line = {
        "metric": "M elem/s assembly; PCG time to 1e-10 (FP32+refine vs FP64); grad/fwd ratio",
        "value": round(value, 2),
        "unit": "M elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kuhn cube, lognormal k, U[-1,1] f)",
        "config": {"workload": "P2 diffusion assembly, 64^3 Kuhn cube jittered (1,572,864 tets), "
                               "11-point rule, lognormal k; step = one FP64 assembly of the full "
                               "matrix into its fixed CSR pattern",
                   "n": args.n, "tets": m.nT, "dofs": ndofs, "nnz_full": nnz,
                   "dofs_interior": nint, "nnz_interior": nnz_int},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_assembly"] if traffic else None,
                     "kernel": "one FP64 assembly = k_asm_rows over the vertex-dof rows + "
                               "k_asm_rows over the edge-dof rows (element factors formed in "
                               "place from each row block's vertex patch in LDS); algorithmic "
                               "bytes %d" % b_asm,
                     "copy_GBps": round(copy_gbs, 1)},
        "assembly_traffic": traffic if traffic else {"note": traffic_note},
        "settle_steps": settle,
        "assembly_f32_ms": round(ms_asm32, 4),
        "symbolic_s": round(t_sym, 3),
        "cpu_baseline": cpu,
    }
    line.update(out_extra)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
