#!/usr/bin/env python3
"""Benchmark: all-mode FP64 direct Helmholtz solve of a 2D x Fourier field.

Workload (one step = one Field::solve over every plane of one field):
16x16 uniform quads on [0,2pi]^2, np=9 GLL, Nz=64 planes (32 modes, the
Nyquist plane skipped -> 63 planes), beta=1, lambda2_k = k^2 + 1/(nu dt)
with nu=0.01, dt=0.005, Dirichlet on all edges, seeded smooth manufactured
solution per plane; direct static condensation + banded Cholesky, FP64.
Inputs are resident in HBM before the timed region.

Prints one JSON line (driver contract).  The dominant kernel's roofline is
measured live with HIP events on the launch stream; the CPU baseline is the
FP64 oracle (test infrastructure) timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Helmholtz DOF-solves/s (all modes) & DNS steps/s at 1e-10 rel err vs FP64"
HBM_PEAK_GBS = 8000.0


def progress(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU-baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip timing the JACPCG / MIXED variants")
    ap.add_argument("--dns-steps", type=int, default=20,
                    help="timed DNS steps (after N_TIME start-up steps); 0 skips the DNS")
    ap.add_argument("--config2", type=int, default=1,
                    help="1: time the 32x32 np=11 Nz=128 MIXED solve in FP64, FP32 and BF16 PCG")
    ap.add_argument("--batch", type=int, default=512,
                    help="independent 2D problems in the batched PCG line; 0 skips it")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 256 MiB (scale for HBM "
                         "byte counters) and print its timing")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (all Fourier modes of a "
              "field resident in one GPU's HBM); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import semtex_amd as sx
    from semtex_amd.workloads import helmholtz_batch, workload_mesh

    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        n = 1 << 25   # 32 Mi doubles = 256 MiB each way
        x = torch.ones(n, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        cms = e0.elapsed_time(e1) / args.steps
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                          "ms": cms, "GBps": 2 * n * 8 / (cms * 1e-3) / 1e9}), flush=True)
        return
    nu, dt, beta = 0.01, 0.005, 1.0
    lam2 = 1.0 / (nu * dt)
    np_, m, nz = workload_mesh("square16_np9_nz64")
    t0 = time.time()
    P = sx.Mesh(np_, m)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, nz // 2, nz, sx.DIRECT)
    t_setup = time.time() - t0
    f, bc, ue = helmholtz_batch(P, nz, lam2, beta, seed=1)
    tf = torch.from_numpy(f.ravel().copy()).to(dev)
    tb = torch.from_numpy(bc.ravel().copy()).to(dev)
    tu = torch.zeros_like(tf)
    nplanes = nz - 1
    npts = P.nglobal + P.nel * P.nint
    dofs = npts * nplanes
    stream = torch.cuda.current_stream()
    L = sx.load()
    import ctypes as C

    def timed(msys, steps, warmup, tol=1e-12):
        """Mean ms per all-plane solve and mean per-kernel ms (second pass
        with HIP events on the launch stream).  Every solve starts from
        u = 0 (the PCG initial guess; the zero fill is inside the timed
        region); the forcing is only read by the solve path."""
        for _ in range(warmup):
            tu.zero_()
            sx.solve(msys, tf, tb, tu, tol_rel=tol, iters=False)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            tu.zero_()
            sx.solve(msys, tf, tb, tu, tol_rel=tol, iters=False)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms_ = ev0.elapsed_time(ev1) / steps
        L.semx_profile_reset()
        L.semx_profile_enable(1)
        for _ in range(steps):
            tu.zero_()
            sx.solve(msys, tf, tb, tu, tol_rel=tol, iters=False)
        torch.cuda.synchronize()
        prof = {}
        for name in ("k_ess_gather", "k_condense", "k_assemble", "k_band_solve",
                     "k_backsub", "k_pcg"):
            tot = C.c_double(); cnt = C.c_longlong()
            L.semx_profile_query(name.encode(), C.byref(tot), C.byref(cnt))
            if cnt.value:
                prof[name] = tot.value / cnt.value
        L.semx_profile_enable(0)
        return ms_, prof

    def graph_timed(msys, steps, tol=1e-12):
        """The same solve captured once into a HIP graph (zero fill + all
        launches) and replayed: mean ms per replay, and whether the replayed
        solution equals the eager one bit for bit."""
        tu.zero_()
        sx.solve(msys, tf, tb, tu, tol_rel=tol, iters=False)
        torch.cuda.synchronize()
        ref = tu.clone()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            tu.zero_()
            sx.solve(msys, tf, tb, tu, tol_rel=tol, iters=False)
        for _ in range(3):
            g.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            g.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        same = bool(torch.equal(tu, ref))
        del g
        return e0.elapsed_time(e1) / steps, same

    def fresh_solve(msys, tol=1e-12):
        tf.copy_(torch.from_numpy(f.ravel().copy()).to(dev))
        tu.zero_()
        it = sx.solve(msys, tf, tb, tu, tol_rel=tol)
        torch.cuda.synchronize()
        return tu.cpu().numpy().reshape(nz, P.planeSize), it

    progress(f"helmholtz setup {t_setup:.2f} s; timing {args.steps} solves")
    ms, kprof = timed(S, args.steps, args.warmup)
    progress(f"direct solve {ms:.4f} ms")
    gms, gsame = graph_timed(S, args.steps)
    progress(f"direct solve as a HIP graph replay {gms:.4f} ms (equal to eager: {gsame})")
    # three fields sharing the system (u, v, w of a DNS step): one call
    fs3 = [tf.clone() for _ in range(3)]
    us3 = [torch.zeros_like(tu) for _ in range(3)]
    for _ in range(3):
        sx.solve_fields(S, fs3, [tb] * 3, us3, iters=False)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        for u3 in us3:
            u3.zero_()
        sx.solve_fields(S, fs3, [tb] * 3, us3, iters=False)
    e1.record(stream)
    torch.cuda.synchronize()
    f3ms = e0.elapsed_time(e1) / args.steps
    f3same = all(bool(torch.equal(u3, us3[0])) for u3 in us3)
    progress(f"three fields in one call {f3ms:.4f} ms")
    u, _ = fresh_solve(S)
    err_exact = float(np.abs(u[:, :P.nPlane] - ue[:, :P.nPlane]).max())

    # algorithmic bytes of k_band_solve per launch: every mode's band factor
    # (nsolve x nband doubles) read once -- the Re and Im planes share it --
    # plus the condensed RHS in and the boundary solution out, per plane;
    # the two-sweep count (the factor streamed by the forward and again by
    # the backward sweep, as dpbtrs does) is reported beside it
    n0 = S.info["nsolve0"]
    ldab = S.info["ldab"]
    bytes_factor = (nz // 2) * n0 * ldab * 8
    bytes_rhs = nplanes * (P.nel * P.next + n0) * 8
    band_bytes = bytes_factor + bytes_rhs
    band_bytes2 = 2 * bytes_factor + bytes_rhs
    band_ms = kprof["k_band_solve"]
    achieved = band_bytes / (band_ms * 1e-3) / 1e9
    # the whole all-mode solve's algorithmic bytes: factor once per mode,
    # one condensed element-matrix set per mode (one geometric set), forcing
    # in and solution out per plane, BC lines
    nline = P.nbound * P.np
    solve_bytes = 8 * ((nz // 2) * (n0 * ldab + P.nsets * (P.next * P.nint + P.nint ** 2))
                       + nplanes * (2 * P.nPlane + nline))

    # HBM bytes per k_band_solve launch from the committed rocprofv3 counter
    # summary (FETCH_SIZE / WRITE_SIZE scaled by a plain copy kernel)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_band_solve_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("traffic_bytes")

    result = {
        "metric": METRIC,
        "value": dofs / (ms * 1e-3),
        "unit": "DOF-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "all-mode direct FP64 Helmholtz: 16x16 quads np=9 on [0,2pi]^2, "
                        "Nz=64 (63 planes), lambda2_k=k^2+1/(nu*dt), nu=0.01, dt=0.005, "
                        "Dirichlet, seeded smooth solutions",
            "dof_per_plane": npts, "planes": nplanes, "nsolve": n0, "nband": ldab,
            "nel": P.nel, "np": P.np, "geometric_sets": P.nsets,
            "setup_s": round(t_setup, 3),
        },
        "roofline": {
            "kernel": "k_band_solve",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": band_bytes,
            "avg_launch_ms": band_ms,
            "frac_two_sweeps": band_bytes2 / (band_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "whole_solve_frac": solve_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "whole_solve_bytes": solve_bytes,
        },
        "kernel_ms": kprof,
        "max_abs_err_vs_exact": err_exact,
        "fields3": {"ms_per_step": f3ms, "value": 3 * npts * nplanes / (f3ms * 1e-3),
                    "fields_equal": f3same,
                    "what": "three fields with the same forcing solved by one solve_fields "
                            "call (Field::solve of u, v, w on one ModalMatrixSys)"},
        "graph": {"ms_per_step": gms, "value": npts * nplanes / (gms * 1e-3),
                  "equals_eager": gsame,
                  "what": "the same solve (zero fill + launches) captured once as a HIP graph "
                          "through the C ABI and replayed"},
    }

    # the iterative methods on the same workload (not the headline)
    variants = {}
    if not args.no_variants:
        for label, method, prec in (("jacpcg_fp64", sx.JACPCG, sx.PCG_FP64),
                                    ("mixed_fp32_refined", sx.MIXED, sx.PCG_FP32_REFINED)):
            V = sx.ModalMatrixSys(P, lam2, beta, 0, nz // 2, nz, method)
            V.set_pcg(prec, inner_tol=1e-4, max_outer=30)
            progress(f"variant {label}")
            uv, it = fresh_solve(V, tol=1e-12)
            vms, vprof = timed(V, max(1, args.steps // 4), max(1, args.warmup // 4))
            variants[label] = {"ms_per_step": vms, "value": dofs / (vms * 1e-3),
                               "iterations_max": it, "kernel_ms": vprof, "u": uv}
            del V

    if not args.no_cpu:
        progress("cpu baseline (oracle)")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        cores = min(16, os.cpu_count() or 1)
        Q = O.Mesh(np_, m)
        fo = np.ascontiguousarray(f.copy())
        uo = np.zeros_like(fo)
        _, t1 = O.field_solve(Q, nz, lam2, beta, 0, fo, bc, uo, nthreads=cores, nrep=1)
        nrep = int(max(1, min(500, args.cpu_seconds / max(t1, 1e-6))))
        fo = np.ascontiguousarray(f.copy())
        uo = np.zeros_like(fo)
        _, tcpu = O.field_solve(Q, nz, lam2, beta, 0, fo, bc, uo, nthreads=cores, nrep=nrep)
        rel = float(np.linalg.norm(u[:, :P.nPlane] - uo[:, :P.nPlane]) /
                    np.linalg.norm(uo[:, :P.nPlane]))
        cpu_model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        result["cpu_baseline"] = {
            "value": dofs / tcpu,
            "unit": "DOF-solves/s",
            "cores": cores,
            "cpu_model": cpu_model,
            "host_threads": os.cpu_count(),
            "kind": "port",
            "sample": f"the same all-mode solve repeated {nrep}x (setup excluded), "
                      f"FP64 oracle, OpenMP over planes; {tcpu * 1e3:.2f} ms per solve",
        }
        result["rel_l2_vs_fp64_oracle"] = rel
        for v in variants.values():
            uv = v["u"]
            v["rel_l2_vs_fp64_oracle"] = float(
                np.linalg.norm(uv[:, :P.nPlane] - uo[:, :P.nPlane]) /
                np.linalg.norm(uo[:, :P.nPlane]))
    for v in variants.values():
        v.pop("u")
    if variants:
        result["iterative_variants"] = variants
    if args.config2:
        result["config2"] = config2_bench(args, torch, sx, L)
    if args.batch > 0:
        result["batch"] = batch_bench(args, torch, sx, L)
    if args.dns_steps > 0:
        result["dns"] = dns_bench(args, torch, sx, L)
    print(json.dumps(result), flush=True)


# inner-iteration bytes per PCG iteration per plane: s (3 nPlane + 17 npts),
# the BLAS-1 and gather/scatter traffic of semtex's loop (src/field.cpp:
# 676-703) with s bytes per vector value (the BF16 operator reads FP32
# vectors, so s = 4 for it too)
PCG_VALUE_BYTES = {"fp64": 8, "fp32_refined": 4, "bf16_refined": 4}


def config2_bench(args, torch, sx, L):
    """The mixed-precision workload: 32x32 quads np=11 on [0,2pi]^2, Nz=128
    (127 planes), lambda2_k = k^2 (beta = 1), Dirichlet, seeded smooth
    solutions; MIXED (mode 0 FP64 DIRECT, modes 1..63 Jacobi-PCG) with the
    PCG in FP64, in FP32 with FP64 refinement, and with the element operator
    on the BF16 matrix cores (FP32 vectors, FP64 refinement), each to
    |r| <= 1e-10 |b|.  The error is measured against the FP64 PCG solution
    at TOL_REL 1e-13."""
    import ctypes as C
    from semtex_amd.workloads import helmholtz_batch, workload_mesh
    dev = torch.device("cuda:0")
    np_, m, nz = workload_mesh("square32_np11_nz128")
    progress("config2 setup (mode 0 factorisation on the host)")
    t0 = time.time()
    P = sx.Mesh(np_, m)
    S = sx.ModalMatrixSys(P, 0.0, 1.0, 0, nz // 2, nz, sx.MIXED)
    t_setup = time.time() - t0
    f, bc, ue = helmholtz_batch(P, nz, 0.0, 1.0, seed=2)
    tf = torch.from_numpy(f.ravel().copy()).to(dev)
    tb = torch.from_numpy(bc.ravel().copy()).to(dev)
    tu = torch.zeros_like(tf)
    npts = P.nglobal + P.nel * P.nint
    nplanes = nz - 1
    stream = torch.cuda.current_stream()
    settings = {"fp64": (sx.PCG_FP64, 1e-5, 1), "fp32_refined": (sx.PCG_FP32_REFINED, 1e-4, 40),
                "bf16_refined": (sx.PCG_BF16_REFINED, 1e-2, 100)}
    S.set_pcg(sx.PCG_FP64)
    tu.zero_()
    sx.solve(S, tf, tb, tu, tol_rel=1e-13, step_max=20000)
    torch.cuda.synchronize()
    uref = tu.clone()
    out = {"workload": "32x32 quads np=11 on [0,2pi]^2, Nz=128 (127 planes), lambda2_k=k^2, "
                       "Dirichlet, seeded smooth solutions; MIXED: mode 0 FP64 DIRECT, modes "
                       "1..63 Jacobi-PCG to |r| <= 1e-10 |b|",
           "dof_per_plane": npts, "planes": nplanes, "setup_s": round(t_setup, 2),
           "bytes_per_iteration_per_plane_formula": "s*(3*nPlane + 17*npts), s = value bytes"}
    for label, (prec, itol, mo) in settings.items():
        S.set_pcg(prec, inner_tol=itol, max_outer=mo)
        progress(f"config2 {label}")
        tu.zero_()
        it = sx.solve(S, tf, tb, tu, tol_rel=1e-10, step_max=20000)
        torch.cuda.synchronize()
        pits = S.plane_iters()
        err = float(torch.linalg.norm(tu - uref) / torch.linalg.norm(uref))
        reps = 3
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            tu.zero_()
            sx.solve(S, tf, tb, tu, tol_rel=1e-10, step_max=20000, iters=False)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        L.semx_profile_reset()
        L.semx_profile_enable(1)
        tu.zero_()
        sx.solve(S, tf, tb, tu, tol_rel=1e-10, step_max=20000, iters=False)
        torch.cuda.synchronize()
        tot = C.c_double(); cnt = C.c_longlong()
        L.semx_profile_query(b"k_pcg", C.byref(tot), C.byref(cnt))
        L.semx_profile_enable(0)
        pcg_ms = tot.value / max(cnt.value, 1)
        sb = PCG_VALUE_BYTES[label]
        byts = float(np.abs(pits).sum()) * sb * (3 * P.nPlane + 17 * npts)
        out[label] = {"ms_per_step": ms, "value": npts * nplanes / (ms * 1e-3),
                      "iterations_max": it, "iterations_sum": int(np.abs(pits).sum()),
                      "rel_l2_vs_fp64_pcg_1e-13": err, "k_pcg_ms": pcg_ms,
                      "k_pcg_inner_bytes": byts,
                      "k_pcg_inner_gb_s": byts / (pcg_ms * 1e-3) / 1e9}
    for label in ("fp32_refined", "bf16_refined"):
        out[label]["speedup_vs_fp64_pcg"] = out["fp64"]["ms_per_step"] / out[label]["ms_per_step"]
    return out


def batch_bench(args, torch, sx, L):
    """Independent 2D problems in one launch: K x K quads, K ~ U{4..16},
    np ~ U{5..11}, lambda2 ~ U[0.1,100], seed 4, Dirichlet, FP32 Jacobi-PCG
    with FP64 refinement to |r| <= 1e-10 |b|; the FP64 PCG batch of the same
    problems (to 1e-13) is the reference for the error."""
    import ctypes as C
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.workloads import batch_problems, single_plane_problem
    dev = torch.device("cuda:0")
    progress(f"batch setup ({args.batch} problems)")
    t0 = time.time()
    meshes, sys64, sys32, sys16, fs, bcs = [], [], [], [], [], []
    dofs = 0
    for (np_, K, lam, cs) in batch_problems(args.batch, seed=4):
        P = sx.Mesh(np_, rect_mesh(K, K, 0, 2 * np.pi, 0, 2 * np.pi))
        f, bc, _ = single_plane_problem(P, lam, cs)
        S64 = sx.ModalMatrixSys(P, lam, 1.0, 0, 1, 1, sx.JACPCG)
        S32 = sx.ModalMatrixSys(P, lam, 1.0, 0, 1, 1, sx.JACPCG)
        S32.set_pcg(sx.PCG_FP32_REFINED, inner_tol=1e-4, max_outer=30)
        S16 = sx.ModalMatrixSys(P, lam, 1.0, 0, 1, 1, sx.JACPCG)
        S16.set_pcg(sx.PCG_BF16_REFINED, inner_tol=1e-2, max_outer=80)
        meshes.append(P); sys64.append(S64); sys32.append(S32); sys16.append(S16)
        fs.append(torch.from_numpy(f.ravel().copy()).to(dev))
        bcs.append(torch.from_numpy(bc.ravel().copy()).to(dev))
        dofs += P.nglobal + P.nel * P.nint
    B64, B32, B16 = sx.Batch(sys64), sx.Batch(sys32), sx.Batch(sys16)
    u64 = [torch.zeros_like(f) for f in fs]
    u32 = [torch.zeros_like(f) for f in fs]
    u16 = [torch.zeros_like(f) for f in fs]
    t_setup = time.time() - t0
    it64 = B64.solve(fs, bcs, u64, tol_rel=1e-13)
    it32 = B32.solve(fs, bcs, u32, tol_rel=1e-10)
    it16 = B16.solve(fs, bcs, u16, tol_rel=1e-10)
    torch.cuda.synchronize()

    def rel_to_fp64(us):
        num = sum(float(((a - b) ** 2).sum().item()) for a, b in zip(us, u64))
        den = sum(float((b ** 2).sum().item()) for b in u64)
        return (num / den) ** 0.5
    rel = rel_to_fp64(u32)
    rel16 = rel_to_fp64(u16)

    def tuner(Bt):
        ops, tms = Bt.info()
        picks = {}
        for P, o in zip(meshes, ops):
            picks.setdefault(o, set()).add((P.np, P.nel))
        return {"tune_ms": round(tms, 1),
                "classes_per_operator": {o: len(c) for o, c in picks.items()}}
    progress(f"batch setup {t_setup:.1f} s; timing")
    stream = torch.cuda.current_stream()

    def timed(Bt, us, tol, reps):
        for _ in range(2):
            for u in us:
                u.zero_()
            Bt.solve(fs, bcs, us, tol_rel=tol, iters=False)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        tot = 0.0
        for _ in range(reps):
            for u in us:
                u.zero_()
            torch.cuda.synchronize()
            e0.record(stream)
            Bt.solve(fs, bcs, us, tol_rel=tol, iters=False)
            e1.record(stream)
            torch.cuda.synchronize()
            tot += e0.elapsed_time(e1)
        return tot / reps

    ms32 = timed(B32, u32, 1e-10, 5)
    ms16 = timed(B16, u16, 1e-10, 5)
    ms64 = timed(B64, u64, 1e-10, 5)
    return {
        "metric": "DOF-solves/s",
        "value": dofs / (ms32 * 1e-3),
        "ms_per_step": ms32,
        "dtype": "f32 inner PCG, f64 refinement",
        "workload": f"{args.batch} independent 2D problems, K x K quads K~U{{4..16}}, "
                    "np~U{5..11}, lambda2~U[0.1,100], seed 4, Dirichlet, Jacobi-PCG to 1e-10, "
                    "one launch (the largest problems on up to 8 workgroups each, the "
                    "rest on one)",
        "dof_total": dofs,
        "iterations_max": it32,
        "fp64_pcg_ms_per_step": ms64,
        "fp64_pcg_value": dofs / (ms64 * 1e-3),
        "rel_l2_vs_fp64": rel,
        "speedup_vs_fp64_pcg": ms64 / ms32,
        "tuner": {"fp64": tuner(B64), "fp32_refined": tuner(B32), "bf16_refined": tuner(B16),
                  "what": "element operator per (np, element count) class chosen by the batch "
                          "tuner on the first solve (block-wide or wavefront-level scalar, or "
                          "BF16 matrix cores)"},
        "bf16_refined": {"ms_per_step": ms16, "value": dofs / (ms16 * 1e-3),
                         "iterations_max": it16, "rel_l2_vs_fp64": rel16,
                         "speedup_vs_fp64_pcg": ms64 / ms16,
                         "what": "the same batch with the inner element operator on the BF16 "
                                 "matrix cores (np padded to 16, FP32 accumulation), FP32 "
                                 "vectors, FP64 refinement"},
        "setup_s": round(t_setup, 2),
    }


def dns_bench(args, torch, sx, L):
    """DNS steps/s: Taylor-Green vortex on [0,2pi]^3 periodic, 24x24 quads
    np=9, Nz=96, Re=1600 (nu=1/1600), dt=1e-3, N_TIME=3, full
    skew-symmetric nonlinear terms, DIRECT velocity and pressure systems;
    timed after the start-up steps, fields resident on the GPU."""
    import ctypes as C
    from semtex_amd.workloads import taylor_green_3d, workload_mesh
    dev = torch.device("cuda:0")
    np_, m, nz = workload_mesh("periodic24_np9_nz96")
    nu, dt, beta, norder = 1.0 / 1600, 1e-3, 1.0, 3
    progress("dns setup (host factorisation of the velocity and pressure systems)")
    t0 = time.time()
    P = sx.Mesh(np_, m)
    D = sx.DNS(P, nz, beta, dt, nu, norder=norder, advection=0, tol_rel=1e-12, step_max=500)
    t_setup = time.time() - t0
    u, v, w = (torch.from_numpy(a.ravel().copy()).to(dev)
               for a in taylor_green_3d(P, nz, beta, seed=3))
    p = torch.zeros_like(u)
    progress(f"dns setup {t_setup:.1f} s; start-up steps")
    D.run(norder + 2, u, v, w, p)          # start-up (JACPCG) steps and warm-up
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    D.run(args.dns_steps, u, v, w, p)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.dns_steps
    L.semx_profile_reset()
    L.semx_profile_enable(1)
    D.run(args.dns_steps, u, v, w, p)
    torch.cuda.synchronize()
    prof = {}
    for name in ("k_band_solve", "k_condense", "k_assemble", "k_backsub", "k_dft_z",
                 "k_grad_xy", "k_grad_nl", "k_grad_z", "k_smooth", "dns_nonlinear"):
        tot = C.c_double(); cnt = C.c_longlong()
        L.semx_profile_query(name.encode(), C.byref(tot), C.byref(cnt))
        if cnt.value:
            prof[name] = {"ms_per_step": tot.value / args.dns_steps,
                          "launches_per_step": cnt.value / args.dns_steps}
    L.semx_profile_enable(0)
    umax = float(u.abs().max().item())
    # the same steps replayed from HIP graphs: one per buffer-role assignment
    # (a cycle of 88 steps here), captured during the untimed warm-up
    D.set_graph(True)
    D.run(100, u, v, w, p)
    torch.cuda.synchronize()
    ng0, _ = D.graph_info()
    e0.record(stream)
    D.run(args.dns_steps, u, v, w, p)
    e1.record(stream)
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / args.dns_steps
    ng1, nrep = D.graph_info()
    D.set_graph(False)
    progress(f"dns eager {ms:.3f} ms/step, graph replay {gms:.3f} ms/step ({ng1} graphs)")
    # algorithmic HBM bytes of the band sweeps, which bound the step: the
    # velocity and the pressure system, every mode's factor read by the
    # forward and the backward sweep (nsolve x nband doubles per mode)
    band_bytes = 2 * (nz // 2) * 2 * P.nsolve * P.nband * 8
    band_ms = prof.get("k_band_solve", {}).get("ms_per_step", float("nan"))
    # the whole step's algorithmic bytes: both systems' factors once per
    # mode, and about 46 compulsory passes over a field (nonlinear terms 6,
    # waveProp 21, setPForce 4, pressure 2, projection 7, velocity 6)
    field_bytes = nz * P.planeSize * 8
    step_bytes = 2 * (nz // 2) * P.nsolve * P.nband * 8 + 46 * field_bytes
    return {
        "metric": "DNS steps/s",
        "value": 1e3 / ms,
        "ms_per_step": ms,
        "steps": args.dns_steps,
        "dtype": "f64",
        "workload": "Taylor-Green + 1e-3 curl perturbation, [0,2pi]^3 periodic, 24x24 quads "
                    "np=9, Nz=96, nu=1/1600, dt=1e-3, N_TIME=3, skew-symmetric, DIRECT",
        "setup_s": round(t_setup, 2),
        "steps_total": D.steps,
        "max_abs_u_fourier": umax,
        "kernel_ms": prof,
        "graph": {"ms_per_step": gms, "value": 1e3 / gms, "graphs": ng1,
                  "captured_in_timed_region": ng1 - ng0,
                  "what": "the same steps replayed from HIP graphs (semx_dns_set_graph), one "
                          "graph per role assignment of the step's buffers"},
        "step_hbm": {"bytes_per_step": step_bytes,
                     "frac_of_peak": step_bytes / (ms * 1e-3) / 8e12,
                     "what": "both systems' band factors once per mode plus 46 field passes "
                             "over the measured step time"},
        "band_hbm": {"bytes_per_step": band_bytes,
                     "achieved_gb_s": band_bytes / (band_ms * 1e-3) / 1e9,
                     "frac_of_peak_in_sweeps": band_bytes / (band_ms * 1e-3) / 8e12,
                     "frac_of_peak_whole_step": band_bytes / (ms * 1e-3) / 8e12,
                     "what": "algorithmic band-factor bytes (2 systems x modes x 2 sweeps x "
                             "nsolve x nband x 8 B) over the sweeps' time and over the whole "
                             "step (field traffic not counted)"},
    }


if __name__ == "__main__":
    main()
