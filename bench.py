"""Headline benchmark: 3D M'4 remesh cycle (P2M + M2P) on one MI355X.

Workload (the metric's configuration): periodic unit cube, 512^3 cells,
134,217,728 particles (one per cell, jittered +-0.3h, advected at CFL 0.5 by
an 8-mode divergence-free field), 3 vorticity components.  One step = one
remesh cycle: P2M of the advected particles (binning included) into the
512^3 node field, then M2P onto a fresh jittered lattice.  Inputs are
generated on the device and resident in HBM before the timed region.

`value` is the plain FP64 form (the reference's double specifics); the
FP32-compensated form is timed in the same run and reported beside it with
its max / rms error against the FP64 GPU result.  The CPU baseline is the
oracle (this repo's C restatement of the reference, OpenMP slab
decomposition over the host cores) on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512, help="cells per axis")
    ap.add_argument("--lda", type=int, default=3)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-f32c", action="store_true", help="skip the FP32-compensated leg")
    ap.add_argument("--cpu-n", type=int, default=256, help="cells per axis of the CPU sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--slab", action="store_true",
                    help="only time one GPU's share of an 8-way z split of the 512^3 cycle: a 512x512x64-cell "
                         "slab with open z faces in the reference's stored box (ghost size 2)")
    ap.add_argument("--config", choices=["c5", "c4"], default="c5",
                    help="c5 (default): the advected lattice; c4: 2^26 particles in 4 Gaussian blobs + uniform, "
                         "P2M then M2P of the same particles")
    ap.add_argument("--no-box", action="store_true", help="skip the stored-box (reference layout) leg")
    ap.add_argument("--paths", action="store_true",
                    help="only time P2M and M2P per path family, tiled against the general path: 3D M'4 and "
                         "BSP2 on the 512^3 advected lattice (M'4 also in a random particle order), 2D M'4 and "
                         "BSP2 on an 8192^2 jittered lattice")
    return ap.parse_args()


def paths_run(core, wl, torch, args):
    """P2M and M2P (FP64, lda 3, binning included) of the tiled kernels against
    the general node-gather / per-particle paths (PPMG_GENERAL=1) on the same
    inputs: 3D BSP2 on the advected 512^3 lattice (M'4 beside it, and M'4 on the
    same particles in a random order), 2D M'4 and
    BSP2 on an 8192^2 lattice jittered by +-0.3h and displaced by up to
    +-0.2h.  Median of three calls after one warm-up, CUDA events around each
    call on the current stream."""
    lda = args.lda

    def one(mesh, x, w, kernel, shape, general):
        if general:
            os.environ["PPMG_GENERAL"] = "1"
        else:
            os.environ.pop("PPMG_GENERAL", None)
        field = torch.empty(shape, dtype=torch.float64, device="cuda")
        out = torch.empty_like(w)
        pp, pm = core.Plan(), core.Plan()
        res = {"p2m": [], "m2p": []}
        for it in range(4):
            for name, call in (("p2m", lambda: core.p2m(mesh, x, w, field, [0] * mesh.dim, kernel, 0, pp)),
                               ("m2p", lambda: core.m2p(mesh, x, field, [0] * mesh.dim, out, kernel, 0, pm))):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                call()
                e1.record()
                torch.cuda.synchronize()
                if it > 0:
                    res[name].append(e0.elapsed_time(e1))
        os.environ.pop("PPMG_GENERAL", None)
        return {k: sorted(v)[1] for k, v in res.items()}

    out = {}
    N = args.n
    mesh, x, w, xf, _ = make_inputs(core, wl, N, lda, args.seed, torch)
    del xf
    for name, kernel in (("3d_mp4", 2), ("3d_bsp2", 1)):
        out[name] = {"cells": f"{N}^3", "particles": N ** 3,
                     "tiled_ms": one(mesh, x, w, kernel, mesh.compact_shape(lda), False),
                     "general_ms": one(mesh, x, w, kernel, mesh.compact_shape(lda), True)}
        print(f"# paths {name}: {out[name]}", file=sys.stderr, flush=True)
    # the same particles in a random order: every lane's particle is then a
    # scattered record and no path reads its stencils in order
    gen = torch.Generator(device="cuda").manual_seed(args.seed + 1)
    order = torch.randperm(N ** 3, generator=gen, device="cuda")
    x, w = x[order].contiguous(), w[order].contiguous()
    del order
    out["3d_mp4_shuffled"] = {"cells": f"{N}^3", "particles": N ** 3,
                              "tiled_ms": one(mesh, x, w, 2, mesh.compact_shape(lda), False),
                              "general_ms": one(mesh, x, w, 2, mesh.compact_shape(lda), True)}
    print(f"# paths 3d_mp4_shuffled: {out['3d_mp4_shuffled']}", file=sys.stderr, flush=True)
    del mesh, x, w
    N2 = 8192 if N == 512 else 16 * N
    n2 = N2 * N2
    mesh2 = core.Mesh(2, N2 + 1)
    gen = torch.Generator(device="cuda").manual_seed(args.seed)
    idx = torch.arange(n2, device="cuda")
    cells = torch.stack([idx % N2, idx // N2], 1).double()
    del idx
    x2 = (cells + 0.5 + 0.6 * (torch.rand((n2, 2), generator=gen, device="cuda", dtype=torch.float64) - 0.5)) / N2
    del cells
    x2 += 0.4 / N2 * (torch.rand((n2, 2), generator=gen, device="cuda", dtype=torch.float64) - 0.5)
    x2 = (x2 - torch.floor(x2)).contiguous()
    w2 = torch.randn((n2, lda), generator=gen, device="cuda", dtype=torch.float64)
    for name, kernel in (("2d_mp4", 2), ("2d_bsp2", 1)):
        out[name] = {"cells": f"{N2}^2", "particles": n2,
                     "tiled_ms": one(mesh2, x2, w2, kernel, mesh2.compact_shape(lda), False),
                     "general_ms": one(mesh2, x2, w2, kernel, mesh2.compact_shape(lda), True)}
        print(f"# paths {name}: {out[name]}", file=sys.stderr, flush=True)
    return {"paths": out, "dtype": "f64", "lda": lda,
            "note": "ms per call, binning included; tiled = the default routing, general = PPMG_GENERAL=1"}


def slab_cycle(core, torch, steps, warmup, lda, seed):
    """P2M + M2P on a 512 x 512 x 64-cell slab (h = 1/512, z faces open),
    16.8M particles on a jittered lattice for each leg.  The field holds the
    unique nodes of the periodic x and y axes (compact) and, along z, the
    reference's 1-g..Nm+g box with g = 2."""
    NX, NY, NZ = 512, 512, 64
    h = 1.0 / NX

    def lattice(s):
        gen = torch.Generator(device="cuda").manual_seed(s)
        n = NX * NY * NZ
        idx = torch.arange(n, device="cuda", dtype=torch.int64)
        cells = torch.stack([idx % NX, (idx // NX) % NY, idx // (NX * NY)], 1).double()
        jit = 0.6 * (torch.rand((n, 3), dtype=torch.float64, device="cuda", generator=gen) - 0.5)
        return ((cells + 0.5 + jit) * h).contiguous()

    x, xq = lattice(seed), lattice(seed + 1)
    u = torch.randn((x.shape[0], lda), dtype=torch.float64, device="cuda")
    mesh = core.Mesh(3, [NX + 1, NY + 1, NZ + 1], [0.0, 0.0, 0.0], [1.0, 1.0, NZ * h], bcdef=[1, 1, 1, 1, 2, 2])
    lo = [0, 0, -2]
    field = torch.empty((NZ + 5, NY, NX, lda), dtype=torch.float64, device="cuda")
    out = torch.empty((xq.shape[0], lda), dtype=torch.float64, device="cuda")
    pp, pm = core.Plan(), core.Plan()
    for _ in range(warmup):
        core.p2m(mesh, x, u, field, lo, 2, core.PREC_F64, pp)
        core.m2p(mesh, xq, field, lo, out, 2, core.PREC_F64, pm)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        core.p2m(mesh, x, u, field, lo, 2, core.PREC_F64, pp)
        core.m2p(mesh, xq, field, lo, out, 2, core.PREC_F64, pm)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return {"metric": "particles/s per GPU, 3D M'4 remesh (P2M+M2P) on one GPU's z-slab of 512^3",
            "value": x.shape[0] / (ms * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic (jittered lattices)",
            "config": {"workload": "512x512x64-cell slab, z faces open, x/y compact, z box with ghost size 2",
                       "particles": x.shape[0], "lda": lda}}


def make_inputs(core, wl, N, lda, seed, torch):
    """Advected jittered lattice + vorticity, and a fresh lattice for M2P (device)."""
    mesh = core.Mesh(3, N + 1)
    x = core.gen_jittered_lattice(N, seed)
    modes = wl.fourier_modes(seed)
    v = core.gen_divfree_velocity(x, modes)
    umax = torch.sqrt((v * v).sum(1)).max().item()
    dt = 0.5 * (1.0 / N) / umax
    x.add_(v, alpha=dt)
    del v
    core.impose_part_bc(mesh, x)
    wmodes = wl.fourier_modes(seed + 1000)
    w = core.gen_divfree_velocity(x, wmodes)                    # vorticity from another 8-mode field
    w /= w.abs().max()
    xfresh = core.gen_jittered_lattice(N, seed + 1)
    return mesh, x, w.contiguous(), xfresh, dt


def clustered_inputs(core, wl, N, n, lda, torch):
    """2^26 particles at 512^3: 80% from 4 Gaussian blobs (sigma 0.05, centres
    from the workload's Philox streams), 20% uniform, wrapped into the unit
    cube, in generation order (not cell order); 3 strengths ~ N(0,1)."""
    import numpy as np
    centres = torch.from_numpy(np.stack([wl.uniforms(4, wl.STREAM_BLOB + d, 4)[0] for d in range(3)], 1)).cuda()
    g = core.gen_normals(n, 3, 44)
    gen = torch.Generator(device="cuda").manual_seed(4)
    uni = torch.rand((n, 3), dtype=torch.float64, device="cuda", generator=gen)
    idx = torch.arange(n, device="cuda")
    x = torch.where((idx < int(0.8 * n))[:, None], centres[idx % 4] + 0.05 * g, uni)
    del g, uni, idx
    x = x - torch.floor(x)
    x = torch.where(x >= 1.0, x - 1.0, x).contiguous()
    return core.Mesh(3, N + 1), x, core.gen_normals(n, lda, 45)


def run_cycle(core, mesh, x, up, field, xfresh, out, prec, p_plan, m_plan, lo=(0, 0, 0), side=None):
    """One remesh cycle: P2M of the advected particles, then M2P onto xfresh.
    With a side stream, the M2P particles (known before the field is) are
    binned there while the P2M deposit runs on the current stream, and the
    gather follows on the side stream once the field is complete."""
    if side is None:
        core.p2m(mesh, x, up, field, list(lo), 2, prec, p_plan)
        core.m2p(mesh, xfresh, field, list(lo), out, 2, prec, m_plan)
        return
    import torch
    main = torch.cuda.current_stream()
    main.wait_stream(side)            # the last gather has read the field
    core.p2m(mesh, x, up, field, list(lo), 2, prec, p_plan)
    with torch.cuda.stream(side):
        core.m2p_prebin(mesh, xfresh, m_plan, 2, prec)
        side.wait_stream(main)        # the field is complete
        core.m2p(mesh, xfresh, field, list(lo), out, 2, prec, m_plan)


def time_form(core, torch, mesh, x, up, xfresh, prec, steps, warmup, ghost=0):
    """FP64 / FP32c cycle on the compact layout (ghost 0) or on the reference's
    stored box 1-g..Nm+g (ghost g)."""
    N = mesh.cells[0]
    lda = up.shape[1]
    vdt = core.VAL_DTYPE[prec]
    lo = (-ghost,) * 3
    shape = mesh.compact_shape(lda) if ghost == 0 else mesh.padded_shape(lda, [ghost] * 3)
    field = torch.empty(shape, dtype=vdt, device="cuda")
    out = torch.empty((xfresh.shape[0], lda), dtype=vdt, device="cuda")
    p_plan, m_plan = core.Plan(), core.Plan()
    p_plan.set_timing(True)
    m_plan.set_timing(True)
    side = torch.cuda.Stream()        # default priority
    for _ in range(warmup):
        run_cycle(core, mesh, x, up, field, xfresh, out, prec, p_plan, m_plan, lo, side)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    t0 = time.perf_counter()
    for _ in range(steps):
        run_cycle(core, mesh, x, up, field, xfresh, out, prec, p_plan, m_plan, lo, side)
    torch.cuda.current_stream().wait_stream(side)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    # Phase durations (HIP events the plans record on their stream) come from
    # as many further cycles run one after another on one stream (reading a
    # plan's timings waits for its stream, which the timed cycles above do
    # not do): the M2P binning is then a phase of its own, where the timed
    # cycles hide it behind the deposit.
    phases = {"bin": [], "weights": [], "deposit": [], "m2p_bin": [], "m2p": []}
    for _ in range(steps):
        run_cycle(core, mesh, x, up, field, xfresh, out, prec, p_plan, m_plan, lo)
        tp = p_plan.timings()
        tm = m_plan.timings()
        phases["bin"].append(tp[0])
        phases["weights"].append(tp[1])
        phases["deposit"].append(tp[2])
        # M2P phases: (binning, gather) on the tiled path, (gather) on the general one
        phases["m2p_bin"].append(tm[0] if len(tm) > 1 else 0.0)
        phases["m2p"].append(tm[-1])
    avg = {k: sum(v) / len(v) for k, v in phases.items()}
    # backward of the last P2M (strength and position gradients of <g, field>),
    # reusing its binning: reported beside the cycle, not part of it
    g = torch.randn(field.shape, dtype=torch.float64, device="cuda").to(vdt)
    core.p2m_adjoint(mesh, x, up, g, list(lo), p_plan, 2, prec)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(3):
        core.p2m_adjoint(mesh, x, up, g, list(lo), p_plan, 2, prec)
    e1.record()
    torch.cuda.synchronize()
    avg["adjoint"] = e0.elapsed_time(e1) / 3
    del g
    return ms, wall, avg, field, out


def clustered_run(core, wl, torch, args):
    """The clustered workload: P2M of 2^26 blob particles at 512^3, then M2P of
    the field at the same particles, FP64 and FP32c; cell occupancy from the
    binning."""
    N, lda = args.n, args.lda
    n = (1 << 26) if N == 512 else N ** 3 // 2
    mesh, x, w = clustered_inputs(core, wl, N, n, lda, torch)
    plan = core.Plan()
    field = torch.empty(mesh.compact_shape(lda), dtype=torch.float64, device="cuda")
    core.p2m(mesh, x, w, field, [0, 0, 0], 2, core.PREC_F64, plan)
    _, _, start = plan.sort_view()
    counts = torch.diff(start.long())
    occ = {"max": int(counts.max().item()), "empty_cell_frac": (counts == 0).double().mean().item(),
           "p99_over_occupied": float(torch.quantile(counts[counts > 0][:1 << 24].double(), 0.99).item())}
    del field, start, counts, plan
    forms = {}
    for name, prec in (("f64", core.PREC_F64), ("f32c", core.PREC_F32C)):
        up = w if prec == core.PREC_F64 else w.float()
        ms, _, ph, f, o = time_form(core, torch, mesh, x, up, x, prec, args.steps, args.warmup)
        forms[name] = {"ms_per_step": ms, "particles_per_s": n / (ms * 1e-3), "phases_ms": ph,
                       "deposit_ns_per_particle": ph["deposit"] * 1e6 / n}
        del f, o
        print(f"# c4 {name} cycle {ms:.3f} ms  phases {ph}", file=sys.stderr, flush=True)
    b64_pass = n * (24 + 8 * lda) + N ** 3 * 8 * lda
    return {"metric": "particles/s 3D M'4 P2M+M2P, clustered particles @512^3",
            "value": forms["f64"]["particles_per_s"], "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": forms["f64"]["ms_per_step"], "higher_is_better": True,
            "dtype": "f64", "data": "synthetic (80% in 4 Gaussian blobs, sigma 0.05; 20% uniform)",
            "config": {"workload": f"clustered P2M+M2P, periodic {N}^3 cells, {n} particles, lda={lda}",
                       "cells": N, "particles": n, "lda": lda},
            "cell_occupancy": occ,
            "cycle_hbm_frac": 2 * b64_pass / (forms["f64"]["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "forms": forms}


def cpu_baseline(oracle_mod, core, wl, torch, n, lda, seed):
    """Oracle P2M + M2P (OpenMP over the host cores of this process, at most
    16: one GPU's share of the host's CPUs) on an n^3 sample."""
    ncores = len(os.sched_getaffinity(0))
    nthreads = min(ncores, 16)
    mesh, x, w, xfresh, _ = make_inputs(core, wl, n, lda, seed, torch)
    xh = x.cpu().numpy()
    uh = w.cpu().numpy()
    xf = xfresh.cpu().numpy()
    del x, w, xfresh
    t = oracle_mod.topo(3, [n + 1] * 3)
    g = [1, 1, 1]
    reps, times = 0, []
    t_start = time.perf_counter()
    while reps < 2 or (reps < 5 and time.perf_counter() - t_start < 20.0):
        t0 = time.perf_counter()
        info, f = oracle_mod.p2m_par(t, 2, xh, uh, g, nthreads)
        info2, back = oracle_mod.m2p_par(t, 2, xf, f, g, nthreads)
        times.append(time.perf_counter() - t0)
        assert info == 0 and info2 == 0
        reps += 1
    best = sorted(times)[len(times) // 2]
    npart = n ** 3
    return {"value": npart / best, "unit": "particles/s", "cores": nthreads, "kind": "port",
            "affinity_cores": ncores,
            "threads_note": "threads = min(affinity cores, 16): one GPU's share of the host is 16 CPUs",
            "sample": f"{n}^3 cells, {npart} particles, lda={lda}, one P2M+M2P remesh cycle (FP64), "
                      f"median of {reps} cycles, {best:.3f} s/cycle"}


def measured_traffic(kernel, N, lda):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (copy-calibrated), when they were taken on this workload."""
    path = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    w = t.get("workload", {})
    k = t.get("kernels", {}).get(kernel)
    if k is None or w.get("cells") != N or w.get("lda") != lda or w.get("dtype") != "f64":
        return None, None
    return k["traffic_bytes"], t.get("source")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding over several GPUs is not part of it)"}))
        sys.exit(2)
    import torch
    from ppm_legacy_amd import core, workloads as wl

    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")     # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8, "launches": 3}))
        return
    if args.slab:
        print(json.dumps(slab_cycle(core, torch, args.steps, args.warmup, args.lda, args.seed)))
        return
    if args.config == "c4":
        print(json.dumps(clustered_run(core, wl, torch, args)))
        return
    if args.paths:
        print(json.dumps(paths_run(core, wl, torch, args)))
        return
    N, lda = args.n, args.lda
    np_ = N ** 3
    mesh, x, w64, xfresh, dt = make_inputs(core, wl, N, lda, args.seed, torch)
    torch.cuda.synchronize()
    print(f"# inputs ready: {N}^3 cells, {np_} particles, lda={lda}, dt={dt:.3e}", file=sys.stderr, flush=True)

    ms64, wall64, ph64, f64, o64 = time_form(core, torch, mesh, x, w64, xfresh, core.PREC_F64, args.steps,
                                             args.warmup)
    print(f"# fp64 cycle {ms64:.3f} ms  phases {ph64}", file=sys.stderr, flush=True)
    # algorithmic bytes (positions FP64 in both forms): per pass Np*(8d + s*lda) + Nn*s*lda
    nn = N ** 3
    b64_pass = np_ * (24 + 8 * lda) + nn * 8 * lda
    res = {}
    if not args.no_f32c:
        w32 = w64.float()
        ms32, wall32, ph32, f32, o32 = time_form(core, torch, mesh, x, w32, xfresh, core.PREC_F32C, args.steps,
                                                 args.warmup)
        fmax = f64.abs().max().item()
        omax = o64.abs().max().item()
        df = (f32.double() - f64)
        do = (o32.double() - o64)
        b32_pass = np_ * (24 + 4 * lda) + nn * 4 * lda
        res = {
            "particles_per_s": np_ / (ms32 * 1e-3),
            "ms_per_step": ms32,
            "speedup_vs_fp64": ms64 / ms32,
            "cycle_hbm_frac": 2 * b32_pass / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "field_err_max_rel": df.abs().max().item() / fmax,
            "field_err_rms_rel": df.pow(2).mean().sqrt().item() / fmax,
            "m2p_err_max_rel": do.abs().max().item() / omax,
            "m2p_err_rms_rel": do.pow(2).mean().sqrt().item() / omax,
            "phases_ms": ph32,
        }
        del f32, o32, w32, df, do
        print(f"# fp32c cycle {ms32:.3f} ms  phases {ph32}", file=sys.stderr, flush=True)

    # the same FP64 cycle on the reference's stored box 1-g..Nm+g (g = 2: the
    # layout the Fortran drop-in passes): node N and the ghost layers are
    # written too, and the adjoint reads its cotangent there
    box = None
    if not args.no_box:
        msb, _, phb, fb, _ob = time_form(core, torch, mesh, x, w64, xfresh, core.PREC_F64, args.steps, args.warmup,
                                         ghost=2)
        uniq = fb[2:2 + N, 2:2 + N, 2:2 + N]
        box = {"ms_per_step": msb, "particles_per_s": np_ / (msb * 1e-3), "ghost": 2,
               "unique_nodes_max_rel_diff_vs_compact": (uniq - f64).abs().max().item() / f64.abs().max().item(),
               "phases_ms": phb}
        del fb, _ob, uniq
        print(f"# fp64 stored-box cycle {msb:.3f} ms  phases {phb}", file=sys.stderr, flush=True)

    # dominant kernel of the FP64 cycle, priced with the algorithmic bytes of its pass
    dom = max(("deposit", "m2p"), key=lambda k: ph64[k])
    dom_kernel = "k_p3_tile" if dom == "deposit" else "k_m3_tile"
    achieved = b64_pass / (ph64[dom] * 1e-3) / 1e9
    # adjoint (strength and position gradients in one launch): read x and u,
    # write dL/du and dL/dx, read the cotangent once
    b_adj = np_ * (24 + 8 * lda + 24 + 8 * lda) + nn * 8 * lda
    traffic, traffic_src = measured_traffic(dom_kernel, N, lda)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        orc.lib()
        orc._setup()
        del f64, o64
        torch.cuda.empty_cache()
        cpu = cpu_baseline(orc, core, wl, torch, args.cpu_n, lda, args.seed)
    out = {
        "metric": "particles/s 3D M'4 remesh (P2M+M2P) @512^3; % HBM peak; FP32c vs FP64 GPU",
        "value": np_ / (ms64 * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox4x32-10 jittered lattice advected by an 8-mode divergence-free field)",
        "config": {"workload": f"3D M'4 remesh cycle (P2M+M2P), periodic {N}^3 cells, {np_} particles, "
                               f"lda={lda}, CFL 0.5 advection, M2P onto a fresh jittered lattice",
                   "cells": N, "particles": np_, "lda": lda, "kernel": "M'4", "seed": args.seed},
        "cycle_hbm_frac": 2 * b64_pass / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "phases_ms": ph64,
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": b64_pass, "kernel_ms": ph64[dom], "traffic_source": traffic_src},
        "adjoint_roofline": {"kernel": "k_a3_tile", "ms": ph64["adjoint"], "algorithmic_bytes": b_adj,
                             "achieved": b_adj / (ph64["adjoint"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": b_adj / (ph64["adjoint"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "fp32c": res or None,
        "stored_box": box,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
