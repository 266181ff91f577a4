"""Benchmark of the MI355X DG operator apply.

Workload (one "step" = one apply over the whole mesh): upwind DG advection,
cell + interior facets, 3D Cartesian 64^3 periodic mesh on [0,1]^3, GLL p=3
(64 DoFs/element, 16,777,216 DoFs), 4 Gauss points/dim, FP64,
beta(x) = (1 + 0.5 sin 2pi y, 0.5 + 0.5 cos 2pi z, 0.25 + 0.25 sin 2pi x),
u = FillRandom(seed) uniform[-1, 1].  Inputs are resident in HBM before the
timed region; three input/output pairs (805 MB) rotate so the 256 MB
Infinity Cache cannot hold the working set.

Prints one JSON line: GDoF/s, ms per apply, the HBM roofline of the apply
kernel (HIP events on the launch stream), the CPU baseline (GenDiL's own
OpenMP apply compiled from its headers, on this host), the FP32 apply, the
6D p-adaptive FP32 apply (8^6 elements, orders 1..3, one launch), and the
FP64 vs FP32-refined GMRES(30) solve of the implicit step (M + dt A) x = M u0
on the 32^3 mesh.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=64, help="elements per axis")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-solve", action="store_true", help="skip the GMRES comparison")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-padapt", action="store_true",
                    help="skip the 6D p-adaptive, product-mesh, adjoint and face-operator entries")
    ap.add_argument("--no-mass", action="store_true", help="skip the 32^3 mass apply")
    ap.add_argument("--no-diffusion", action="store_true", help="skip the 64^3 SIPDG diffusion apply")
    ap.add_argument("--no-general", action="store_true",
                    help="skip the 64^3 applies with fields the Kronecker kernels do not take")
    ap.add_argument("--general-only", action="store_true",
                    help="only time the 64^3 applies with non-Kronecker fields (one JSON object)")
    ap.add_argument("--face-ops-only", action="store_true",
                    help="only time the face-centric operators (one JSON object)")
    ap.add_argument("--product6d-only", action="store_true",
                    help="only time GenDiL's 6D benchmark operator (5^6 and 10^6 elements; one JSON object)")
    ap.add_argument("--padapt-classes", action="store_true",
                    help="only time the 6D p-adaptive apply on the mixed mesh and on one mesh per order (one JSON object)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also time a plain device copy of the apply's input/output bytes (counter calibration)")
    return ap.parse_args()


def measured_traffic(key):
    """HBM bytes per launch of the default apply kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain
    copy); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k3k_stream_adv3d_64cubed_p3_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


PRIME_SECONDS = 0.1


def time_applies(op, xs, ys, steps, warmup):
    """ms per apply over `steps` applies on the current stream (HIP events),
    after `warmup` untimed applies.  Before those, applies run untimed for
    PRIME_SECONDS of wall time: from idle the MI355X takes some 30 ms of load
    to reach its working clock (three back-to-back 200-apply timings of the
    64^3 FP64 apply read 77, 69 and 67 us per apply)."""
    import time

    import torch

    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    i = 0
    while time.perf_counter() - t0 < PRIME_SECONDS:
        op(xs[i % len(xs)], ys[i % len(ys)])
        i += 1
        if i % 32 == 0:
            torch.cuda.synchronize()
    for i in range(warmup):
        op(xs[i % len(xs)], ys[i % len(ys)])
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for i in range(steps):
        op(xs[i % len(xs)], ys[i % len(ys)])
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms per apply


def host_cpu():
    """CPU model, physical cores of this host and of this process's CPU set,
    and the OpenMP threads the CPU baseline uses: the CPU share this process
    is given (OMP_NUM_THREADS where the environment sets it, else the
    physical cores of its CPU set)."""
    model, phys_all, phys_mine = "unknown", set(), set()
    try:
        mine = os.sched_getaffinity(0)
    except AttributeError:
        mine = set(range(os.cpu_count() or 1))
    try:
        cpu = pid = cid = None
        with open("/proc/cpuinfo") as f:
            for line in f.read().splitlines() + [""]:
                if not line.strip():
                    if cpu is not None:
                        phys_all.add((pid, cid))
                        if cpu in mine:
                            phys_mine.add((pid, cid))
                    cpu = pid = cid = None
                    continue
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    cpu = int(v)
                elif k == "model name":
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    cid = v
    except OSError:
        pass
    cores_mine = len(phys_mine) or len(mine)
    env = os.environ.get("OMP_NUM_THREADS")
    threads = int(env) if env and env.isdigit() and int(env) > 0 else cores_mine
    return {"model": model, "host_physical_cores": len(phys_all) or None, "process_cpus": len(mine),
            "process_physical_cores": cores_mine, "threads": threads}


def cpu_baseline(n, pairs=5):
    """GenDiL's own OpenMP apply (SerialKernelConfiguration, compiled from its
    headers into oracle/_ref/gendil_ref) on the 64^3 workload, timed with its
    TimeOperator pattern (benchmark-common.hpp:276-295): one warm-up pair,
    then `pairs` timed pairs of applies."""
    import oracle

    cpu = host_cpu()
    threads = cpu["threads"]
    env = {"OMP_NUM_THREADS": str(threads), "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}
    if oracle.ref_available():
        r = oracle.ref_bench("adv", 3, 3, 4, "L", n, True, "c2", "none", 12345, pairs, env=env)
        return {
            "value": r["dofs_per_s"] / 1e9,
            "unit": "GDoF/s",
            "cores": threads,
            "kind": "reference",
            "cpu": cpu,
            "sample": f"GenDiL SerialKernelConfiguration apply (OpenMP, {threads} threads, OMP_PROC_BIND=close, "
            f"OMP_PLACES=cores) compiled from its headers, full {n}^3 p=3 workload, 1 warm-up + {pairs} timed "
            f"apply pairs",
        }
    # the oracle's restatement where GenDiL's binary is absent, on a 32^3 sample
    import numpy as np

    import gendil_amd as G
    from gendil_amd import workloads as W

    m = min(n, 32)
    mesh = G.CartesianMesh.unit(3, m, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    P = oracle.Problem(fes, G.MakeIntegrationRule(4), adv_coef=1.0, adv=W.advection_c2(), num_threads=threads)
    x = G.fill_random(P.num_dofs, 12345)
    P.apply(x)
    t0 = time.perf_counter()
    for _ in range(2 * pairs):
        P.apply(x)
    t = (time.perf_counter() - t0) / (2 * pairs)
    return {
        "value": P.num_dofs / t / 1e9,
        "unit": "GDoF/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu,
        "sample": f"CPU restatement (oracle) of GenDiL's apply, {m}^3 p=3 workload, {2 * pairs} timed applies",
    }


def mass_32(steps, warmup, n=32, pairs=10):
    """The mass apply y = M x on the 32^3 mesh (GLL p=3, 4 Gauss points,
    sigma = 1, 2.1M DoFs, FP64): `pairs` input/output pairs rotate (16.8 MB
    each, 336 MB in all), so the 256 MB Infinity Cache cannot hold the working
    set; algorithmic bytes 16 per DoF.  Beside it GenDiL's own OpenMP mass
    apply where it is built."""
    import torch

    import gendil_amd as G

    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, n, False), G.MakeLobattoFiniteElement(3))
    op = G.MakeMassFiniteElementOperator(fes, G.MakeIntegrationRule(4))
    N = op.num_dofs
    x0 = torch.as_tensor(G.fill_random(N, 12345), device="cuda")
    xs = [x0 * (1.0 - 0.01 * i) for i in range(pairs)]
    ys = [torch.empty_like(x0) for _ in range(pairs)]
    ms = time_applies(op, xs, ys, steps, warmup)
    t = ms * 1e-3
    out = {
        "workload": f"mass apply, 3D {n}^3, GLL p=3, 4 GL pts/dim, sigma = 1, FP64, {pairs} rotating vector pairs",
        "dofs": N,
        "value": N / t / 1e9,
        "unit": "GDoF/s",
        "ms_per_step": ms,
        "variant": op.plan.query_variant(torch.float64, 1.0, 0.0),
        "roofline": {"bound": "hbm", "achieved": 16.0 * N / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": 16.0 * N / t / 1e9 / HBM_PEAK_GBS},
    }
    del xs, ys
    try:
        import oracle

        if oracle.ref_available():
            threads = host_cpu()["threads"]
            r = oracle.ref_bench("mass", 3, 3, 4, "L", n, False, "one", "none", 12345, 5,
                                 env={"OMP_NUM_THREADS": str(threads), "OMP_PROC_BIND": "close",
                                      "OMP_PLACES": "cores"})
            out["cpu_reference"] = {"value": r["dofs_per_s"] / 1e9, "unit": "GDoF/s", "cores": threads,
                                    "kind": "reference", "sample": "1 warm-up + 5 timed apply pairs"}
    except Exception as e:  # the GPU figure stands on its own
        out["cpu_reference"] = {"error": str(e)[:200]}
    return out


def diffusion_64(steps, warmup, n=64, q=5):
    """The SIPDG diffusion apply (GenDiL's DiffusionOperator) on the 64^3
    periodic mesh, GLL p=3, q = p + 2 Gauss points (its Poisson examples'
    rule), diffusivity 2 + 0.25 sin 2 pi x + 0.17 cos 2 pi (2y + 0.13) +
    0.11 sin 2 pi (3z + 0.37) + 0.07 sin 2 pi (x + 2y + 3z), sigma = -1,
    kappa = 16; FP64, three rotating vector pairs (402 MB each way).
    Algorithmic bytes 16 per DoF.  Beside it GenDiL's own OpenMP apply on the
    32^3 mesh where it is built."""
    import torch

    import gendil_amd as G
    from gendil_amd import workloads as W

    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, n, True), G.MakeLobattoFiniteElement(3))
    op = G.MakeDiffusionOperator(fes, G.MakeIntegrationRule(q), W.diffusivity_trig_3d(), -1.0, 16.0)
    N = op.num_dofs
    x0 = torch.as_tensor(G.fill_random(N, 12345), device="cuda")
    xs = [x0, x0 * 0.5, x0 * -0.25]
    ys = [torch.empty_like(x0) for _ in range(3)]
    ms = time_applies(op, xs, ys, steps, warmup)
    t = ms * 1e-3
    out = {
        "workload": f"SIPDG diffusion apply, 3D {n}^3 periodic, GLL p=3, {q} GL pts/dim, trig diffusivity, "
                    "sigma = -1, kappa = 16, FP64",
        "dofs": N,
        "value": N / t / 1e9,
        "unit": "GDoF/s",
        "ms_per_step": ms,
        "roofline": {"bound": "hbm", "achieved": 16.0 * N / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": 16.0 * N / t / 1e9 / HBM_PEAK_GBS},
    }
    del xs, ys
    try:
        import oracle

        if oracle.ref_available():
            threads = host_cpu()["threads"]
            r = oracle.ref_bench("diff", 3, 3, q, "L", 32, True, "trig", "s-1_k16", 12345, 2,
                                 env={"OMP_NUM_THREADS": str(threads), "OMP_PROC_BIND": "close",
                                      "OMP_PLACES": "cores"})
            out["cpu_reference"] = {"value": r["dofs_per_s"] / 1e9, "unit": "GDoF/s", "cores": threads,
                                    "kind": "reference", "sample": "32^3 mesh, 1 warm-up + 2 timed apply pairs"}
    except Exception as e:  # the GPU figure stands on its own
        out["cpu_reference"] = {"error": str(e)[:200]}
    return out


def solve_comparison():
    import torch

    import gendil_amd as G
    from gendil_amd import solvers as S
    from gendil_amd import workloads as W

    n = 32
    h = 1.0 / n
    mesh = G.CartesianMesh.unit(3, n, False)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    dt = 0.5 * h / W.advection_c2_max_norm()
    K = G.MakeMassAdvectionOperator(fes, ir, W.advection_c2(), bc=0.0, mass_coef=1.0, adv_coef=dt)
    M = G.MakeMassFiniteElementOperator(fes, ir)
    N = K.num_dofs
    u0 = W.gaussian_bump(fes, (0.5, 0.5, 0.5), 0.1) + 1e-3 * G.fill_random(N, 2024)
    u0 = torch.as_tensor(u0, device="cuda")
    b = torch.empty_like(u0)
    M(u0, b)
    out = {}
    for name, fn in (("fp64", S.gmres), ("mixed", S.gmres_mixed)):
        runs = []
        for _ in range(3):
            x = torch.zeros_like(b)
            torch.cuda.synchronize()
            res = fn(K, b, x, restart=30, rel_tol=1e-10)
            runs.append((res.seconds, res, x))
        best = min(runs, key=lambda r: r[0])
        out[name] = best
    xr = out["fp64"][2]
    xm = out["mixed"][2]
    return {
        "workload": "(M + dt A) x = M u0, 32^3 p=3, inflow g=0, dt = 0.5 h/|beta|max, GMRES(30) to 1e-10",
        "fp64_seconds": out["fp64"][0],
        "mixed_seconds": out["mixed"][0],
        "ratio_mixed_over_fp64": out["mixed"][0] / out["fp64"][0],
        "fp64_iterations": out["fp64"][1].iterations,
        "mixed_iterations": out["mixed"][1].iterations,
        "mixed_refinements": out["mixed"][1].refinements,
        "fp64_final_rel_residual": out["fp64"][1].final_rel_residual,
        "mixed_final_rel_residual": out["mixed"][1].final_rel_residual,
        "mixed_vs_fp64_solution_rel_diff": float((xm - xr).norm() / xr.norm()),
    }


FP32_PEAK_TFLOPS = 157.3  # MI355X published dense FP32 (counts packed dual FMAs)
FP32_UNPACKED_TFLOPS = 78.6  # 256 CUs x 64 lanes x 2 flop x 2.4 GHz: one FMA per lane per clock


def kronecker_6d_fma(fes, beta_active, beta_neg):
    """Algorithmic FMA count of one Kronecker 6D apply (gd_pa6k.hpp): per
    element of N nodes/dim, N^4 slices x 2 N^3 FMA per pair operator, with
    MM applied four times and one G operator per active component; per inflow
    face from an NB-node neighbour, five sweeps of N^(k+1) NB^(5-k) FMA."""
    import numpy as np

    m = fes.mesh
    sizes = np.asarray(m.sizes)
    E = m.num_cells
    n = fes.elem_order.astype(np.int64) + 1
    subs = np.array(np.unravel_index(np.arange(E), sizes[::-1]))[::-1]  # x fastest
    pairs = [(2, 5), (1, 4), (0, 3)]
    ncomp = sum(int(beta_active[a]) + int(beta_active[b]) for a, b in pairs)
    total = float(np.sum(2.0 * n.astype(np.float64) ** 7 * (4 + ncomp)))
    strides = np.cumprod(np.concatenate(([1], sizes[:-1])))
    for c in range(6):
        if not beta_active[c]:
            continue
        step = 1 if beta_neg[c] else -1  # inflow neighbour
        s = subs[c] + step
        ok = (s >= 0) & (s < sizes[c])
        if m.periodic[c]:
            s = s % sizes[c]
            ok[:] = True
        nb = np.arange(E) + (s - subs[c]) * strides[c]
        nbn = np.where(ok, n[np.where(ok, nb, 0)], 0).astype(np.float64)
        N = n.astype(np.float64)
        face = sum(N ** (k + 1) * nbn ** (5 - k) for k in range(5))
        total += float(np.sum(np.where(ok, face, 0.0)))
    return total


def padapt_6d(steps, warmup, n=8):
    """6D phase-space advection, n^6 elements, per-element orders p in {1,2,3}
    (probabilities 0.5/0.35/0.15, seed 2024), q = p + 1, FP32, one launch per
    apply; the thread layout comes from the autotuner."""
    import torch

    import gendil_amd as G
    from gendil_amd import workloads as W

    fes, adv = W.phase_space_6d_padaptive(n, seed=2024)
    op = G.MakeAdvectionOperator(fes, None, adv)
    N = op.num_dofs
    E = fes.mesh.num_cells
    chosen, tuned_ms = op.plan.autotune(torch.float32, 0.0, 1.0, reps=3)
    gen = torch.Generator(device="cuda").manual_seed(2024)
    x0 = torch.rand(N, generator=gen, device="cuda", dtype=torch.float32) * 2.0 - 1.0
    xs = [x0, x0 * 0.5, x0 * -0.25]
    ys = [torch.empty_like(x0) for _ in range(3)]
    ms = time_applies(op, xs, ys, steps, warmup)
    t = ms * 1e-3
    # algorithmic bytes: u read + y written (FP32), per element its order
    # (1 B), DoF offset (8 B) and bucket entry (4 B)
    nbytes = 8.0 * N + 13.0 * E
    y = ys[(steps - 1) % 3].double()
    cons = float(y.sum().abs() / y.abs().sum())
    counts = [int((fes.elem_order == p).sum()) for p in (1, 2, 3)]
    # error of the FP32 apply against the FP64 apply of the same input
    xlast = xs[(steps - 1) % 3]
    del xs, ys
    y64 = torch.empty(N, dtype=torch.float64, device="cuda")
    op(xlast.double(), y64)
    err = float((y - y64).norm() / y64.norm())
    del y64, y, xlast
    variant = op.plan.query_variant(torch.float32, 0.0, 1.0)
    compute = None
    if variant == 2:
        # beta = (X3, X4, X5, 0.1, 0, 0): components 0..3 active, all >= 0
        fma = kronecker_6d_fma(fes, [True, True, True, True, False, False], [False] * 6)
        tflops = 2.0 * fma / t / 1e12
        compute = {
            "bound": "valu_fp32",
            "achieved": tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": tflops / FP32_PEAK_TFLOPS,
            "frac_of_unpacked_fma_rate": tflops / FP32_UNPACKED_TFLOPS,
            "algorithmic_gflop_per_apply": 2.0 * fma / 1e9,
        }
    return {
        "workload": f"6D phase space {n}^6 (space periodic), p in {{1,2,3}} w.p. 0.5/0.35/0.15, q = p+1, "
        "beta = (X3, X4, X5, 0.1, 0, 0), FP32, one launch",
        "dofs": N,
        "elements_per_order": {"p1": counts[0], "p2": counts[1], "p3": counts[2]},
        "value": N / t / 1e9,
        "unit": "GDoF/s",
        "ms_per_step": ms,
        "roofline_frac": nbytes / t / 1e9 / HBM_PEAK_GBS,
        "compute_roofline": compute,
        "autotuned_variant": chosen,
        "variant": variant,
        "conservation_sum_over_abs_sum": cons,
        "err_vs_fp64": err,
    }


def padapt_classes(steps, warmup, n=8):
    """The 6D p-adaptive Kronecker apply (variant 2) on the mixed-order 8^6
    mesh and on three 8^6 meshes of one order each (p = 1, 2, 3): per-element
    cost of each order class, and its share of the mixed apply."""
    import numpy as np
    import torch

    import gendil_amd as G
    from gendil_amd import workloads as W
    from gendil_amd.discretization import MakePAdaptiveSpace

    mesh = G.CartesianMesh([n] * 6, [1.0 / n] * 6, [0.0] * 6, [True, True, True, False, False, False])
    mixed = W.padaptive_orders(mesh.num_cells, 2024)
    out = {}
    for name, orders in [("mixed", mixed)] + [(f"p{p}", np.full(mesh.num_cells, p, dtype=mixed.dtype)) for p in (1, 2, 3)]:
        fes = MakePAdaptiveSpace(mesh, orders, orders + 1)
        op = G.MakeAdvectionOperator(fes, None, W.phase_space_6d_affine())
        op.plan.set_variant(torch.float32, 2)
        N = op.num_dofs
        gen = torch.Generator(device="cuda").manual_seed(1)
        x0 = torch.rand(N, generator=gen, device="cuda", dtype=torch.float32)
        xs = [x0, x0 * 0.5, x0 * -0.25]
        ys = [torch.empty_like(x0) for _ in range(3)]
        ms = time_applies(op, xs, ys, steps, warmup)
        out[name] = {"ms": ms, "dofs": N, "ns_per_element": ms * 1e6 / mesh.num_cells}
        del xs, ys, x0, op, fes
        torch.cuda.empty_cache()
    counts = {f"p{p}": int((mixed == p).sum()) for p in (1, 2, 3)}
    est = {k: out[k]["ns_per_element"] * counts[k] * 1e-6 for k in counts}
    total = sum(est.values())
    out["mixed_elements_per_order"] = counts
    out["estimated_share_of_mixed"] = {k: v / total for k, v in est.items()}
    out["estimated_mixed_ms"] = total
    return out


def product_6d(steps, warmup, n=5):
    """GenDiL's own 6D benchmark operator (benchmarks/benchmark-advection-6d.cpp):
    3D x 2D x 1D product mesh, n elements per axis (h = 1, 1.2345, 1), Legendre
    p=1, 3 Gauss points per axis, beta = 1, no boundary data, FP64; beside it
    the reference's own CPU apply of the same operator where it is built."""
    import torch

    import gendil_amd as G
    from gendil_amd.fields import SeparableField, VectorField

    mesh = G.MakeCartesianProductMesh(G.Cartesian3DMesh(1.0, n, n, n), G.Cartesian2DMesh(1.2345, n, n),
                                      G.Cartesian1DMesh(1.0, n))
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLegendreFiniteElement(1))
    op = G.MakeAdvectionOperator(fes, G.MakeIntegrationRule(3), VectorField([SeparableField.constant(1.0)] * 6))
    N = op.num_dofs
    x0 = torch.as_tensor(G.fill_random(N, 12345), device="cuda")
    xs = [x0, x0 * 0.5, x0 * -0.25]
    ys = [torch.empty_like(x0) for _ in range(3)]
    ms = time_applies(op, xs, ys, steps, warmup)
    out = {
        "workload": f"GenDiL 6D benchmark operator: product mesh {n}^6 (3D x 2D x 1D), Legendre p=1, q=3, beta=1, FP64",
        "dofs": N,
        "value": N / (ms * 1e-3) / 1e9,
        "unit": "GDoF/s",
        "ms_per_step": ms,
        "roofline_frac": 16.0 * N / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    try:
        import oracle

        if oracle.ref_available():
            threads = host_cpu()["threads"]
            r = oracle.ref_bench("adv6dprod", 6, 1, 3, "G", n, False, "ones", "none", 12345, 5,
                                 env={"OMP_NUM_THREADS": str(threads), "OMP_PROC_BIND": "close",
                                      "OMP_PLACES": "cores"})
            out["cpu_reference"] = {"value": r["dofs_per_s"] / 1e9, "unit": "GDoF/s", "cores": threads,
                                    "kind": "reference"}
    except Exception as e:  # the GPU figure stands on its own
        out["cpu_reference"] = {"error": str(e)[:200]}
    out["variant"] = op.plan.query_variant(torch.float64, 0.0, 1.0)
    del xs, ys, x0, op
    torch.cuda.empty_cache()
    # the same operator on 10^6 elements (64 M DoFs, 512 MB per vector): the
    # HBM roofline of the kernel, which the 1 M-DoF benchmark size is too
    # small to reach
    big = 10
    mesh = G.MakeCartesianProductMesh(G.Cartesian3DMesh(1.0, big, big, big), G.Cartesian2DMesh(1.2345, big, big),
                                      G.Cartesian1DMesh(1.0, big))
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLegendreFiniteElement(1))
    op = G.MakeAdvectionOperator(fes, G.MakeIntegrationRule(3), VectorField([SeparableField.constant(1.0)] * 6))
    N = op.num_dofs
    x0 = torch.rand(N, dtype=torch.float64, device="cuda")
    xs = [x0, x0 * 0.5, x0 * -0.25]
    ys = [torch.empty_like(x0) for _ in range(3)]
    ms = time_applies(op, xs, ys, max(5, steps // 4), 2)
    out["large"] = {
        "workload": f"same operator, product mesh {big}^6",
        "dofs": N,
        "value": N / (ms * 1e-3) / 1e9,
        "unit": "GDoF/s",
        "ms_per_step": ms,
        "roofline_frac": 16.0 * N / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    del xs, ys, x0
    return out


def adjoint_48(steps, warmup, n=48, p=4, q=5):
    """The adjoint workload: 3D 48^3 periodic, GLL p=4, 5 Gauss points, FP64,
    beta per quadrature point (benchmark field + 0.1 N(0,1), seed 7): the
    forward apply, A^T r and the beta contraction, each timed alone.
    Algorithmic bytes: the vectors each kernel reads and writes once plus the
    per-point field (3 q^3 cell values per element, q^2 per face)."""
    import torch

    import gendil_amd as G
    from gendil_amd import workloads as W

    mesh = G.CartesianMesh.unit(3, n, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(p))
    ir = G.MakeIntegrationRule(q)
    cell, faces = W.noisy_pointwise_beta(mesh, q, 7)
    bc = torch.as_tensor(cell, device="cuda")
    bf = [torch.as_tensor(f, device="cuda") for f in faces]
    del cell, faces
    op = G.MakeAdvectionOperator(fes, ir, G.PointwiseField(bc, bf))
    plan = op.plan
    N = op.num_dofs
    gen = torch.Generator(device="cuda").manual_seed(7)
    u = torch.rand(N, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    r = torch.rand(N, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    y = torch.empty_like(u)
    dcell = torch.empty_like(bc)
    dfaces = [torch.empty_like(f) for f in bf]
    beta_bytes = 8.0 * (bc.numel() + sum(f.numel() for f in bf))

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    res = {"workload": f"adjoint: 3D {n}^3 periodic, GLL p={p}, {q} GL points, per-point beta, FP64", "dofs": N}
    for name, fn, nbytes in (
        ("forward", lambda: plan.apply(0.0, 1.0, u, y), 16.0 * N + beta_bytes),
        ("transpose", lambda: plan.apply_transpose(0.0, 1.0, r, y), 16.0 * N + beta_bytes),
        ("grad_beta", lambda: plan.grad_beta(1.0, u, r, dcell, dfaces), 16.0 * N + beta_bytes),
    ):
        ms = timed(fn)
        res[name] = {"ms": ms, "GB_per_s": nbytes / (ms * 1e-3) / 1e9,
                     "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    del bc, bf, dcell, dfaces, u, r, y
    return res


def general_64(steps, warmup, n=64):
    """The C2 mesh and basis (3D 64^3 periodic, GLL p=3, 4 Gauss points,
    FP64) with two fields the Kronecker kernels do not take:
      rotating: beta = (0.5 - y, x - 0.5, 0.25), each component a function of
        one coordinate but changing sign inside the mesh (the register
        kernel, with its per-point upwind choice);
      two-coordinate: beta = (1 + 0.5 sin 2pi y cos 2pi z,
        0.5 + 0.5 cos 2pi z sin 2pi x, 0.25 + 0.25 sin 2pi x cos 2pi y), each
        component a product over two coordinates (the generic kernel).
    Algorithmic bytes: u read and y written once, 16 B per DoF."""
    import numpy as np
    import torch

    import gendil_amd as G
    from gendil_amd.fields import SeparableField, VectorField

    tp = 2.0 * math.pi
    fields = {
        "rotating": VectorField([
            SeparableField([(1.0, {1: lambda y: 0.5 - y})]),
            SeparableField([(1.0, {0: lambda x: x - 0.5})]),
            SeparableField.constant(0.25),
        ]),
        "two_coordinate": VectorField([
            SeparableField([(1.0, {}), (0.5, {1: lambda y: np.sin(tp * y), 2: lambda z: np.cos(tp * z)})]),
            SeparableField([(0.5, {}), (0.5, {2: lambda z: np.cos(tp * z), 0: lambda x: np.sin(tp * x)})]),
            SeparableField([(0.25, {}), (0.25, {0: lambda x: np.sin(tp * x), 1: lambda y: np.cos(tp * y)})]),
        ]),
    }
    mesh = G.CartesianMesh.unit(3, n, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    out = {"workload": f"3D {n}^3 periodic, GLL p=3, 4 GL pts/dim, FP64, fields the Kronecker kernels do not take"}
    x0 = None
    for name, beta in fields.items():
        op = G.MakeAdvectionOperator(fes, ir, beta)
        N = op.num_dofs
        if x0 is None:
            x0 = torch.as_tensor(G.fill_random(N, 12345), device="cuda")
            xs = [x0, x0 * 0.5, x0 * -0.25]
            ys = [torch.empty_like(x0) for _ in range(3)]
        ms = time_applies(op, xs, ys, steps, warmup)
        out[name] = {
            "value": N / (ms * 1e-3) / 1e9,
            "unit": "GDoF/s",
            "ms_per_step": ms,
            "roofline_frac": 16.0 * N / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "variant": op.plan.query_variant(torch.float64, 0.0, 1.0),
        }
        del op
    del xs, ys, x0
    return out


def face_ops(steps, warmup, n=64, p=3):
    """GenDiL's face-centric advection operators on device-resident vectors:
    the interior faces of a 64^3 non-periodic mesh (GLL p=3, 4 Gauss points,
    the benchmark field) and the +x interface of two 64^3 meshes of orders 1
    and 3 sharing 5 points (beta = (y, z, x), the p-adaptivity test's field).
    Both accumulate into y, so the algorithmic bytes are x read, y read and y
    written once: 24 B per DoF."""
    import numpy as np
    import torch

    import gendil_amd as G
    from gendil_amd import workloads as W

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    out = {}
    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, n, False), G.MakeLobattoFiniteElement(p))
    op = G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity(fes.mesh.sizes),
                                     G.MakeIntegrationRule(p + 1), W.advection_c2())
    N = fes.num_dofs
    gen = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(N, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    y = torch.zeros_like(x)
    ms = timed(lambda: op(x, y))
    out["interior"] = {
        "workload": f"interior faces, 3D {n}^3 non-periodic, GLL p={p}, {p + 1} GL points, benchmark field, FP64",
        "dofs": N, "ms": ms, "value": N / (ms * 1e-3) / 1e9, "unit": "GDoF/s",
        "roofline_frac": 24.0 * N / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    del x, y
    try:
        import oracle

        if oracle.ref_available():
            threads = host_cpu()["threads"]
            r = oracle.ref_bench("advface", 3, p, p + 1, "L", n, False, "c2", "none", 12345, 5,
                                 env={"OMP_NUM_THREADS": str(threads), "OMP_PROC_BIND": "close",
                                      "OMP_PLACES": "cores"})
            out["interior"]["cpu_reference"] = {"value": r["dofs_per_s"] / 1e9, "unit": "GDoF/s", "cores": threads,
                                                "kind": "reference"}
    except Exception as e:  # the GPU figure stands on its own
        out["interior"]["cpu_reference"] = {"error": str(e)[:200]}
    h = 1.0 / n
    fm = G.MakeFiniteElementSpace(G.Cartesian3DMesh(h, n, n, n), G.MakeLobattoFiniteElement(1))
    fp = G.MakeFiniteElementSpace(G.Cartesian3DMesh(h, n, n, n, origin=(1.0, 0.0, 0.0)),
                                  G.MakeLobattoFiniteElement(3), G.L2Restriction(fm.num_dofs))
    yzx = G.AffineField(np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]), np.zeros(3))
    im = G.MakeAdvectionFaceOperator(fm, fp, G.CartesianIntermeshFaceConnectivity(3, 3, (n, n, n), (n, n, n)),
                                     G.MakeIntegrationRule(5), yzx)
    M = fm.num_dofs + fp.num_dofs
    x = torch.rand(M, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    y = torch.zeros_like(x)
    ms = timed(lambda: im(x, y))
    faces = n * n
    out["intermesh"] = {
        "workload": f"+x interface of two {n}^3 meshes, GLL p=1 | p=3, 5 shared GL points, beta=(y,z,x), FP64",
        "faces": faces, "ms": ms, "faces_per_s": faces / (ms * 1e-3),
    }
    del x, y
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(
            json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it", "n_gpus": args.gpus})
        )
        sys.exit(2)
    import torch

    import gendil_amd as G
    from gendil_amd import workloads as W

    if args.face_ops_only:
        print(json.dumps({"face_ops": face_ops(args.steps, args.warmup)}))
        return
    if args.product6d_only:
        print(json.dumps({"product6d": product_6d(args.steps, args.warmup)}))
        return
    if args.general_only:
        print(json.dumps({"general64": general_64(args.steps, args.warmup)}))
        return
    if args.padapt_classes:
        print(json.dumps({"padapt_classes": padapt_classes(args.steps, args.warmup)}))
        return
    n = args.n
    mesh = G.CartesianMesh.unit(3, n, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    op = G.MakeAdvectionOperator(fes, ir, W.advection_c2())
    N = op.num_dofs
    x0 = torch.as_tensor(G.fill_random(N, 12345), device="cuda")
    xs = [x0, x0 * 0.5, x0 * -0.25]
    ys = [torch.empty_like(x0) for _ in range(3)]
    ms = time_applies(op, xs, ys, args.steps, args.warmup)
    t = ms * 1e-3
    bytes_per_apply = 16.0 * N  # read u once, write y once (FP64)
    achieved = bytes_per_apply / t / 1e9
    result = {
        "metric": "operator apply GDoF/s + % HBM roofline; FP32-refined vs FP64 solve time @1e-10",
        "value": N / t / 1e9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (FillRandom-identical mt19937 inputs, seed 12345)",
        "timing": f"HIP events around the K applies on the launch stream, after {PRIME_SECONDS * 1e3:.0f} ms of untimed "
                  "applies (clock ramp) and the W warm-up applies; three input/output pairs rotate",
        "config": {
            "workload": f"upwind DG advection apply, 3D {n}^3 periodic, GLL p=3, 4 GL pts/dim, C2 beta field",
            "dofs": N,
            "elements": n ** 3,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic("traffic_bytes_per_launch_f64") if n == 64 else None,
        },
    }
    if args.copy_calibration:
        # a plain copy moving the apply's algorithmic bytes (read N, write N
        # values): hardware byte counters of the apply are scaled by what the
        # same counters read for this kernel
        for i in range(args.steps):
            ys[i % 3].copy_(xs[i % 3])
        torch.cuda.synchronize()
        result["copy_calibration_bytes"] = bytes_per_apply
    del ys
    if not args.no_fp32:
        x32 = [x.float() for x in xs]
        y32 = [torch.empty_like(x32[0]) for _ in range(3)]
        ms32 = time_applies(op, x32, y32, args.steps, args.warmup)
        # error of the FP32 apply against the FP64 apply of the same input
        y64 = torch.empty_like(xs[0])
        op(x32[0].double(), y64)
        op(x32[0], y32[0])
        err32 = float((y32[0].double() - y64).norm() / y64.norm())
        result["fp32"] = {
            "value": N / (ms32 * 1e-3) / 1e9,
            "ms_per_step": ms32,
            "roofline_frac": 8.0 * N / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "err_vs_fp64": err32,
        }
        del x32, y32, y64
    del xs
    torch.cuda.empty_cache()
    if not args.no_mass:
        result["mass32"] = mass_32(args.steps, args.warmup)
        torch.cuda.empty_cache()
    if not args.no_general:
        result["general64"] = general_64(max(10, args.steps // 4), 3)
        torch.cuda.empty_cache()
    if not args.no_diffusion:
        result["diffusion64"] = diffusion_64(max(10, args.steps // 4), 3)
        torch.cuda.empty_cache()
    if not args.no_padapt:
        result["padapt6d"] = padapt_6d(max(10, args.steps // 4), 3)
        torch.cuda.empty_cache()
        result["product6d"] = product_6d(max(20, args.steps // 2), 5)
        torch.cuda.empty_cache()
        result["adjoint48"] = adjoint_48(max(10, args.steps // 10), 2)
        torch.cuda.empty_cache()
        result["face_ops"] = face_ops(max(10, args.steps // 10), 2)
        torch.cuda.empty_cache()
    if not args.no_solve:
        result["solve"] = solve_comparison()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
