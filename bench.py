#!/usr/bin/env python3
"""Benchmark: forward + adjoint MLUPS of the D3Q19 MRT design-gradient pass on one MI355X.

Workload (the headline configuration): D3Q19 incompressible MRT with Brinkman design
coupling, 384 x 192 x 192 duct (Hecht velocity inlet u=0.04, Hecht pressure outlet rho=1,
bounce-back side walls), objective = pressure drop, T = 3000 time steps, 48 HBM checkpoint
slots placed by binomial checkpointing for the known horizon (the fewest recomputations; the
reference's Wang dynamic placement is timed beside it on the same solver and gives the same
gradient bit for bit). Synthetic design field: SplitMix64 uniform, seed 0, 5x5x5 box filter.

One bench step = one design-gradient evaluation: equilibrium init, T recorded forward
sweeps, T reverse sweeps with checkpoint recomputation. MLUPS counts every grid node once
per time step:  value = nodes * T / (time of one step) / 1e6, FP64 plain form.
The FP16-shifted-storage form is measured the same way and reported under "forms".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lbm-topopt_amd"))

METRIC = "forward+adjoint MLUPS and % peak HBM BW, D3Q19 MRT, FP64 plain vs FP16-storage"
PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bytes_per_lu(form, Q=19):
    """Algorithmic bytes per lattice update (forward, adjoint)."""
    sf, sl, sg = {"fp64": (8, 8, 8), "fp32": (4, 4, 4), "fp16": (2, 4, 4), "bf16": (2, 4, 4)}[form]
    return 2 * Q * sf + sg, Q * sf + 2 * Q * sl + sg + 2 * 8


def run_form(lbmx, workloads, form, nx, ny, nz, horizon, slots, steps, warmup, sweeps, mode, compare=None):
    prec = {"fp64": lbmx.FP64, "fp16": lbmx.FP16_SHIFTED, "bf16": lbmx.BF16_SHIFTED, "fp32": lbmx.FP32}[form]
    d = lbmx.make_desc(nx, ny, nz, lattice=lbmx.D3Q19, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN,
                       precision=prec, omega=1.6, interp_q=0.1, alpha_max=0.1)
    s = lbmx.Solver(d)
    workloads.duct3d(s, nx, ny, nz, u_in=0.04, time_const=1.0 / horizon)
    gamma = workloads.design_field(nx, ny, nz, seed=0)
    s.set_design(gamma)
    s.checkpoint_config(slots, mode)
    N = nx * ny * nz
    # kernel-only sweep timing (HIP events on the engine's stream)
    s.init_equilibrium()
    s.time_sweeps(0, 5)
    fwd_ms = s.time_sweeps(0, sweeps) / sweeps
    s.time_sweeps(1, 3)
    adj_ms = s.time_sweeps(1, sweeps) / sweeps
    log(f"{form}: forward sweep {fwd_ms:.4f} ms, adjoint sweep {adj_ms:.4f} ms")
    for w in range(warmup):
        t0 = time.perf_counter()
        s.gradient(horizon)
        log(f"{form}: warmup step {w} {time.perf_counter() - t0:.2f} s")
    times, fphase, rphase = [], [], []
    J = None
    recompute = []
    for k in range(steps):
        before = s.stats()["recompute_sweeps"]
        t0 = time.perf_counter()
        J, sens = s.gradient(horizon, want_sens=(k == steps - 1))
        dt = time.perf_counter() - t0
        st = s.stats()
        recompute.append(st["recompute_sweeps"] - before)
        times.append(dt)
        fphase.append(st["forward_phase_ms"] / 1e3)
        rphase.append(st["reverse_phase_ms"] / 1e3)
        log(f"{form}: step {k} {dt:.2f} s (forward {fphase[-1]:.2f} s, reverse {rphase[-1]:.2f} s, "
            f"recompute sweeps {recompute[-1]})")
    other = None
    if compare is not None:
        # the same gradient with another checkpoint placement: same states, same sweeps
        name, cmode = compare
        s.checkpoint_config(slots, cmode)
        before = s.stats()["recompute_sweeps"]
        t0 = time.perf_counter()
        J2, sens2 = s.gradient(horizon)
        dt = time.perf_counter() - t0
        other = {"placement": name, "ms_per_step": 1e3 * dt, "mlups_fwd_adj": N * horizon / dt / 1e6,
                 "recompute_sweeps_per_gradient": s.stats()["recompute_sweeps"] - before,
                 "bitwise_equal": bool(J2 == J and np.array_equal(sens2, sens))}
        log(f"{form}: {name} placement {dt:.2f} s, recompute sweeps {other['recompute_sweeps_per_gradient']}, "
            f"bitwise equal {other['bitwise_equal']}")
    st = s.stats()
    dev_bytes = s.device_bytes()
    s.close()
    t_step = max(times)
    bf, ba = bytes_per_lu(form)
    res = {
        "ms_per_step": 1e3 * t_step,
        "mlups_fwd_adj": N * horizon / t_step / 1e6,
        "mlups_forward_phase": N * horizon / max(fphase) / 1e6,
        "mlups_reverse_phase": N * horizon / max(rphase) / 1e6,
        "fwd_sweep_ms": fwd_ms,
        "adj_sweep_ms": adj_ms,
        "mlups_fwd_kernel": N / (fwd_ms * 1e3),
        "mlups_adj_kernel": N / (adj_ms * 1e3),
        "fwd_gbs": bf * N / (fwd_ms * 1e-3) / 1e9,
        "adj_gbs": ba * N / (adj_ms * 1e-3) / 1e9,
        "fwd_bytes_per_lu": bf,
        "adj_bytes_per_lu": ba,
        "objective": J,
        "_sens": sens,
        "recompute_sweeps_per_gradient": recompute[-1],
        # algorithmic bytes of the whole pass (recorded and recomputed forward sweeps, adjoint
        # sweeps) over its wall time, host hand-off and initialisation included
        "pass_gbs": ((horizon + recompute[-1]) * bf + horizon * ba) * N / t_step / 1e9,
        "device_gb": dev_bytes / 1e9,
    }
    if other is not None:
        res["other_placement"] = other
    return res


def run_ensemble(lbmx, workloads, steps, slots, warmup, mode, placement, more_slots=None):
    """Robust ensemble (24 members: 3 projection realisations x 8 candidates on uneven grids,
    PBB<IncompressibleCascaded2d> at tau = 0.6, FP32): batched gradient vs the same 24
    gradients as sequential single-instance runs. Aggregate MLUPS = total nodes x steps / s."""
    def mk(nx, ny):
        return lbmx.Solver(lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.CASCADED,
                                          design=lbmx.DESIGN_PBB, precision=lbmx.FP32, omega=1.0 / 0.6,
                                          interp_q=0.1, alpha_max=1.0))
    members = [s for s, _, _ in workloads.ensemble_members(mk, steps=steps)]
    nodes = sum(m.nodes for m in members)
    b = lbmx.Batch(members)
    b.checkpoint_config(slots, mode)
    rows, tune_ms = b.autotune(10)
    paired = b.tile()[1]
    log(f"ensemble: {len(members)} members, {nodes} nodes, autotuned tile rows {rows}{' paired' if paired else ''} "
        f"(ms per fwd+adj sweep for rows 1/2/4: {', '.join(f'{v:.3f}' for v in tune_ms)})")
    for _ in range(warmup):
        b.gradient(steps, want_sens=False)
    t0 = time.perf_counter()
    Jb, _ = b.gradient(steps)
    tb = time.perf_counter() - t0
    st = b.stats()
    log(f"ensemble: batched gradient {tb:.2f} s (forward {st['forward_phase_ms'] / 1e3:.2f} s, "
        f"reverse {st['reverse_phase_ms'] / 1e3:.2f} s)")
    b.set_graph(True)
    b.gradient(steps, want_sens=False)  # captures the pass into a HIP graph
    t0 = time.perf_counter()
    Jg, _ = b.gradient(steps)
    tg = time.perf_counter() - t0
    b.set_graph(False)
    log(f"ensemble: batched gradient replayed from a HIP graph {tg:.2f} s")
    for m in members:
        m.checkpoint_config(slots, mode)
    for _ in range(warmup):
        for m in members:
            m.gradient(steps, want_sens=False)
    t0 = time.perf_counter()
    Js = [m.gradient(steps)[0] for m in members]
    ts = time.perf_counter() - t0
    log(f"ensemble: sequential single-instance gradients {ts:.2f} s")
    Js = np.array(Js)
    if not (np.all(np.isfinite(Js)) and np.all(np.isfinite(Jb)) and np.all(np.isfinite(Jg))):
        raise RuntimeError(f"ensemble objective not finite: batched {Jb}, single {Js}")
    jerr = float(max(np.max(np.abs(Jb - Js) / np.abs(Js)), np.max(np.abs(Jg - Js) / np.abs(Js))))
    dev = b.device_bytes()
    # the same batched pass with more of HBM given to checkpoints (fewer recomputed sweeps)
    more = {}
    if more_slots and more_slots > slots:
        b.checkpoint_config(more_slots, mode)
        b.gradient(steps, want_sens=False)
        t0 = time.perf_counter()
        Jm, _ = b.gradient(steps)
        tm = time.perf_counter() - t0
        if not np.array_equal(Jm, Jb):
            raise RuntimeError("ensemble objective depends on the checkpoint count")
        more = {"checkpoints": more_slots, "batched_s": round(tm, 3),
                "batched_mlups": round(nodes * steps / tm / 1e6, 1),
                "recompute_sweeps": int(lbmx.checkpoint_simulate(more_slots, steps, mode)[0]),
                "batch_device_gb": round(b.device_bytes() / 1e9, 2)}
        log(f"ensemble: batched gradient with {more_slots} checkpoints {tm:.2f} s")
    b.close()
    for m in members:
        m.close()
    return {"members": len(members), "nodes": nodes, "time_steps": steps, "checkpoints": slots,
            "placement": placement,
            "tile_rows": rows, "tile_forward_paired": paired,
            "autotune_ms_rows_1_2_4": [round(v, 4) for v in tune_ms],
            "batched_s": round(tb, 3), "batched_graph_s": round(tg, 3), "sequential_s": round(ts, 3),
            "batched_mlups": round(nodes * steps / tb / 1e6, 1),
            "batched_graph_mlups": round(nodes * steps / tg / 1e6, 1),
            "sequential_mlups": round(nodes * steps / ts / 1e6, 1),
            "speedup": round(ts / tb, 3), "max_rel_J_diff_batched_vs_single": jerr,
            "batch_device_gb": round(dev / 1e9, 2), "more_checkpoints": more}


def cone_taps(R, dims):
    """Lattice offsets d with |d| <= R (the cone's taps per filtered node)."""
    r = np.arange(-R, R + 1)
    if dims == 2:
        return int(np.sum(r[:, None] ** 2 + r[None, :] ** 2 <= R * R))
    return int(np.sum(r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2 <= R * R))


def run_design_pipeline(lbmx, reps=20, cpu=True):
    """Design pipeline (mapping, density filter, projection, and the sensitivity chain) on the
    device: the two-stage chain of applications/RobustPump2d/robustPump2d.cc:140-146
    (radii 8 / 4, beta 2 / 1) on the 1024 x 512 grid, and one projecting stage of radius 3 on
    the 384 x 192 x 192 grid. Device time per call from HIP events around its kernels."""
    B = lbmx.make_box
    cases = [("2d_double_projection_1024x512", (1024, 512, 1), (64, 959, 1, 510, 0, 0),
              ((8, True, 2.0, 0.5), (4, True, 1.0, 0.5))),
             ("3d_projection_384x192x192", (384, 192, 192), (32, 351, 1, 190, 1, 190), ((3, True, 2.0, 0.5),))]
    out = {}
    rng = np.random.default_rng(0)
    for name, dims, dbox, stages in cases:
        f = lbmx.Filter(lbmx.make_filter_desc(*dims, B(*dbox), B(*dbox), stages))
        d = rng.random(f.n_design)
        s = rng.standard_normal(f.n_full)
        f.design(d)
        f.sensitivities(s)
        tf, tb = [], []
        for _ in range(reps):
            f.design(d)
            a, _ = f.timing()
            f.sensitivities(s)
            _, b = f.timing()
            tf.append(a)
            tb.append(b)
        f.close()
        n = dims[0] * dims[1] * dims[2]
        taps = sum(cone_taps(st[0], 3 if dims[2] > 1 else 2) for st in stages)
        ms_f, ms_b = float(np.median(tf)), float(np.median(tb))
        # algorithmic HBM bytes per node: forward map 24, per stage cone 24 (in, den, out) and
        # projection 16; backward per stage scale 32 (sens, x, den, out) and cone 16, extraction 16
        bf = 24 + sum(24 + (16 if st[1] else 0) for st in stages)
        bb = sum(32 + 16 for _ in stages) + 16
        out[name] = {"nodes": n, "stages": [list(st) for st in stages], "cone_taps_per_node": taps,
                     "design_ms": round(ms_f, 4), "sensitivities_ms": round(ms_b, 4),
                     "gnodes_per_s": round(n / ((ms_f + ms_b) * 1e-3) / 1e9, 3),
                     "design_gbs": round(bf * n / (ms_f * 1e-3) / 1e9, 1),
                     "sensitivities_gbs": round(bb * n / (ms_b * 1e-3) / 1e9, 1),
                     "cone_gflops": round(2 * 2 * taps * n / ((ms_f + ms_b) * 1e-3) / 1e9, 1)}
        log(f"design pipeline {name}: design {ms_f:.3f} ms, sensitivities {ms_b:.3f} ms")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # noqa: E402  (baseline leg only)
        nx, ny = 256, 128
        ch = oracle.FilterChain(nx, ny, 1, (16, 239, 1, 126, 0, 0), stages=cases[0][3])
        d = rng.random(ch.nd)
        t0 = time.perf_counter()
        for _ in range(3):
            ch.design(d)
            ch.sensitivities(np.ones(ch.n))
        dt = (time.perf_counter() - t0) / 3
        out["cpu_baseline"] = {"value": round(nx * ny / dt / 1e9, 6), "unit": "Gnodes/s (design + sensitivities)",
                               "cores": 1, "kind": "port",
                               "sample": f"oracle chain (radii 8 / 4) on {nx}x{ny}, assembled rows, one thread"}
        log(f"design pipeline cpu baseline: {dt * 1e3:.1f} ms per design + sensitivities on {nx}x{ny}")
    return out


def run_mma(lbmx, cpu=True):
    """Method of moving asymptotes on the device: the robust form of
    applications/RobustPump2d/robustPump2d.cc:173-190 (three realisations with a_j = 1 and a
    volume constraint with a_j = 0) on synthetic smooth responses, at the 1024 x 512 and the
    384 x 192 x 192 design sizes. Wall milliseconds per update (host round trips of the dual
    Newton iteration included), with host arrays and with device arrays (lbmx_mma_update_device);
    the oracle's single-thread update as CPU baseline."""
    out = {}

    def inputs(x, targets, weights):
        n = x.size
        g = [w * np.sum((x - t) ** 2) / n for t, w in zip(targets, weights)] + [x.mean() - 0.45]
        dg = [2 * w * (x - t) / n for t, w in zip(targets, weights)] + [np.full(n, 1.0 / n)]
        return np.zeros(n), np.array(g), np.stack(dg)

    import torch
    rng = np.random.default_rng(0)
    for name, n in (("1024x512", 524288), ("384x192x192", 14155776)):
        targets = [0.2 + 0.6 * rng.random(n) for _ in range(3)]
        w = [1.0, 1.3, 0.8]
        h = lbmx.MMA(n, 4, a=[1.0, 1.0, 1.0, 0.0])
        hd = lbmx.MMA(n, 4, a=[1.0, 1.0, 1.0, 0.0])   # the same updates on device arrays
        x = np.full(n, 0.5)
        xd = torch.from_numpy(x).to("cuda")
        ms, its, msd, same = [], [], [], True
        for _ in range(4):
            dfdx, g, dg = inputs(x, targets, w)
            dfdx_d = torch.from_numpy(dfdx).to("cuda")
            dg_d = [torch.from_numpy(r).to("cuda") for r in dg]
            x = h.update(x, dfdx, g, dg, 0.1)
            st = h.state()
            ms.append(st["ms"])
            its.append(st["newton_iters"])
            hd.update_device(xd, dfdx_d, g, dg_d, 0.1)
            msd.append(hd.state()["ms"])
            same = same and bool(np.array_equal(xd.cpu().numpy(), x))
            del dfdx_d, dg_d
        h.close()
        hd.close()
        out[name] = {"n": n, "m": 4, "update_ms": [round(v, 3) for v in ms], "newton_iters": its,
                     "ms_per_newton_iter": round(float(np.median(np.array(ms) / np.array(its))), 4),
                     "device_arrays_update_ms": [round(v, 3) for v in msd], "device_arrays_bitwise_equal": same}
        log(f"mma {name}: {', '.join(f'{v:.1f}' for v in ms)} ms per update ({its} dual Newton iterations); "
            f"device arrays {', '.join(f'{v:.1f}' for v in msd)} ms, bitwise equal {same}")
        del targets, xd
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # noqa: E402  (baseline leg only)
        n = 131072
        targets = [0.2 + 0.6 * np.random.default_rng(0).random(n) for _ in range(3)]
        ref = oracle.MMA(n, 4, a=[1.0, 1.0, 1.0, 0.0])
        x = np.full(n, 0.5)
        t0 = time.perf_counter()
        its = 0
        for _ in range(2):
            x = ref.update(x, *inputs(x, targets, [1.0, 1.3, 0.8]), 0.1)
            its += ref.state()["newton_iters"]
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(dt * 1e3, 1), "unit": "ms for two updates", "cores": 1, "kind": "port",
                               "newton_iters": its,
                               "ns_per_variable_newton_iter": round(dt * 1e9 / (n * its), 4),
                               "sample": f"oracle MMA, n = {n}, m = 4, first two updates from x = 0.5"}
        log(f"mma cpu baseline: {dt * 1e3:.1f} ms for two updates ({its} Newton iterations), n = {n}")
    for name in ("1024x512", "384x192x192"):
        r = out[name]
        r["ns_per_variable_newton_iter"] = round(r["ms_per_newton_iter"] * 1e6 / r["n"], 4)
    return out


def read_traffic(key):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated by the
    copy kernel) for the 384x192x192 workload, from profiles/traffic_d3q19_mrt_brinkman.txt."""
    path = os.path.join(ROOT, "profiles", "traffic_d3q19_mrt_brinkman.txt")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            k, _, v = line.strip().partition("=")
            if k == key:
                return float(v)
    return None


# The workloads of the reference's benchmark configurations besides the headline duct:
# (name, lattice, collision, grid, tau or omega, time steps, checkpoint slots, setup, initial state).
def other_workloads(lbmx, workloads):
    def ch256(obj, nx, ny, steps):
        workloads.channel2d(obj, nx, ny, time_const=1.0 / steps)

    def ch1024(obj, nx, ny, steps):
        workloads.modulated_channel(obj, nx, ny, steps=steps)

    def th2048(obj, nx, ny, steps):
        workloads.thermal_channel2d(obj, nx, ny, time_const=1.0 / steps)

    bgk = dict(lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_BRINKMAN, omega=1.0 / 0.8, interp_q=0.1,
               alpha_max=0.1)
    th = dict(lattice=lbmx.D2Q9_D2Q5, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, omega=1.25, interp_q=0.1,
              alpha_max=0.1, thermal_diffusivity=0.02, ramp_p=1.0, ramp_ck=0.1)
    return [
        {"name": "channel_256x128_bgk_brinkman_forward", "grid": (256, 128), "desc": bgk, "steps": 2000, "slots": None,
         "setup": ch256, "init": {}, "Q": 9},
        {"name": "channel_1024x512_bgk_brinkman_modulated", "grid": (1024, 512), "desc": bgk, "steps": 5000,
         "slots": 64, "setup": ch1024, "init": {}, "Q": 9},
        {"name": "thermal_2048x1024_mrt", "grid": (2048, 1024), "desc": th, "steps": 4000, "slots": 64,
         "setup": th2048, "init": {"u0": (0.03, 0.0, 0.0), "T0": 0.5}, "Q": 14},
    ]


def run_other_workloads(lbmx, workloads, mode, sweeps=50, only=None):
    """Forward and adjoint sweep rates (HIP events) and the whole pass of the other benchmark
    workloads, FP64 and FP16-shifted forms. The 256 x 128 channel is a forward run; the others
    are design gradients with checkpoints. Their grids are small enough to be partly
    L2-resident, so their GB/s are algorithmic bytes over time, not HBM traffic."""
    out = {}
    for w in other_workloads(lbmx, workloads):
        if only and w["name"] not in only:
            continue
        nx, ny = w["grid"]
        N = nx * ny
        rec = {}
        for form, prec in (("fp64", lbmx.FP64), ("fp16", lbmx.FP16_SHIFTED)):
            s = lbmx.Solver(lbmx.make_desc(nx, ny, precision=prec, **w["desc"]))
            w["setup"](s, nx, ny, w["steps"])
            s.set_design(workloads.design_field(nx, ny, 1, seed=0))
            bf, ba = bytes_per_lu(form, Q=w["Q"])
            s.init_equilibrium(**w["init"])
            s.time_sweeps(0, 5)
            fms = s.time_sweeps(0, sweeps) / sweeps
            r = {"fwd_sweep_ms": round(fms, 5), "mlups_fwd_kernel": round(N / (fms * 1e3), 1),
                 "fwd_gbs": round(bf * N / (fms * 1e-3) / 1e9, 1), "fwd_bytes_per_lu": bf}
            if w["slots"] is None:
                s.init_equilibrium(**w["init"])
                t0 = time.perf_counter()
                J = s.forward(w["steps"])
                dt = time.perf_counter() - t0
                r.update({"forward_run_s": round(dt, 4), "mlups_forward_run": round(N * w["steps"] / dt / 1e6, 1),
                          "objective": J})
            else:
                s.time_sweeps(1, 3)
                ams = s.time_sweeps(1, sweeps) / sweeps
                r.update({"adj_sweep_ms": round(ams, 5), "mlups_adj_kernel": round(N / (ams * 1e3), 1),
                          "adj_gbs": round(ba * N / (ams * 1e-3) / 1e9, 1), "adj_bytes_per_lu": ba})
                s.checkpoint_config(w["slots"], mode)
                s.gradient(w["steps"], want_sens=False, **w["init"])  # warm-up
                t0 = time.perf_counter()
                J, g = s.gradient(w["steps"], **w["init"])
                dt = time.perf_counter() - t0
                st = s.stats()
                r.update({"gradient_s": round(dt, 4), "mlups_fwd_adj": round(N * w["steps"] / dt / 1e6, 1),
                          "mlups_forward_phase": round(N * w["steps"] / (st["forward_phase_ms"] / 1e3) / 1e6, 1),
                          "mlups_reverse_phase": round(N * w["steps"] / (st["reverse_phase_ms"] / 1e3) / 1e6, 1),
                          "objective": J, "_sens": g})
            s.close()
            rec[form] = r
            log(f"{w['name']} {form}: forward sweep {fms:.4f} ms"
                + (f", adjoint sweep {r['adj_sweep_ms']:.4f} ms, gradient {r['gradient_s']:.3f} s"
                   if "gradient_s" in r else f", {w['steps']}-step run {r['forward_run_s']:.3f} s"))
        if "_sens" in rec["fp64"]:
            g64, g16 = rec["fp64"].pop("_sens"), rec["fp16"].pop("_sens")
            J64, J16 = rec["fp64"]["objective"], rec["fp16"]["objective"]
            rec["fp16"]["vs_fp64"] = {"objective_rel_err": abs(J16 - J64) / abs(J64),
                                      "gradient_rel_err_max_norm": float(np.max(np.abs(g16 - g64)) / np.max(np.abs(g64))),
                                      "mlups_ratio": round(rec["fp16"]["mlups_fwd_adj"] / rec["fp64"]["mlups_fwd_adj"], 3)}
        out[w["name"]] = {"nodes": N, "time_steps": w["steps"], "checkpoints": w["slots"], **rec}
    return out


def _host_cores():
    """(logical CPUs, physical cores per socket, sockets) from /proc/cpuinfo."""
    logical, cores, sockets = 0, None, set()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "processor":
                    logical += 1
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
                elif k == "physical id":
                    sockets.add(v.strip())
    except OSError:
        pass
    return logical, cores, max(1, len(sockets))


def reference_adjoint_report(lbmx, workloads, oracle, threads, steps=3000, ncells=8):
    """The reference's own adjoint of the cascaded collision, run as it is coded (the CPU
    oracle's as-coded variant of ReverseIncompressibleCascaded2d), on one member of the robust
    ensemble at its full horizon (256 x 128, eta 0.5, PBB<cascaded> at tau 0.6, FP64), against
    the engine's gradient (the true transpose) and central differences of J on the GPU."""
    nx, ny = 256, 128
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.CASCADED, design=lbmx.DESIGN_PBB, omega=1.0 / 0.6,
                       interp_q=0.1, alpha_max=1.0)
    g = workloads.heaviside_project(workloads.ensemble_candidate(nx, ny, 1000), 8.0, 0.5)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, time_const=1.0 / steps)
    s.set_design(g)
    s.checkpoint_config(64, lbmx.CHECKPOINT_BINOMIAL)
    J, sens = s.gradient(steps)
    o = oracle.Oracle(d, threads=threads)
    workloads.channel2d(o, nx, ny, time_const=1.0 / steps)
    o.set_design(g)
    o.set_reference_reverse(True)
    t0 = time.perf_counter()
    _, sref, _ = o.gradient(steps)
    t_cpu = time.perf_counter() - t0
    del o
    scale = float(np.max(np.abs(sens)))
    cells = np.concatenate([np.argsort(-np.abs(sref - sens))[:ncells // 2],
                            np.random.default_rng(0).choice(g.size, ncells - ncells // 2, replace=False)])
    h = 1e-5
    fd = []
    for c in cells:
        v = []
        for sgn in (1, -1):
            gg = g.copy()
            gg[c] += sgn * h
            s.set_design(gg)
            s.init_equilibrium()
            v.append(s.forward(steps))
        fd.append((v[0] - v[1]) / (2 * h))
    s.close()
    fd = np.array(fd)
    r = {"member": "256x128 channel, eta 0.5, PBB<IncompressibleCascaded2d> tau 0.6, FP64", "steps": steps,
         "reference_vs_true_rel_max": float(np.max(np.abs(sref - sens)) / scale),
         "true_vs_central_fd_rel_max": float(np.max(np.abs(sens[cells] - fd)) / scale),
         "reference_vs_central_fd_rel_max": float(np.max(np.abs(sref[cells] - fd)) / scale),
         "fd_cells": int(len(cells)), "oracle_gradient_s": round(t_cpu, 2)}
    log(f"reference adjoint on an ensemble member: vs true {r['reference_vs_true_rel_max']:.2e}; vs central FD: "
        f"true {r['true_vs_central_fd_rel_max']:.1e}, reference {r['reference_vs_central_fd_rel_max']:.1e}")
    return r


def cpu_baseline(workloads, lbmx, horizon_head=3000, slots_head=48, mode_name="binomial"):
    """The CPU oracle (the reference's algorithm restated, FP64, OpenMP over nodes) on the
    host's cores, on every benchmark workload at its full grid and a reduced horizon: forward
    MLUPS from forward steps, reverse-phase MLUPS from the reverse steps of a gradient whose
    states are all kept, and the design-gradient rate of the full horizon with the engine's
    checkpoint schedule (same slots and placement, its recomputation counted from the schedule's
    host replay), modelled from those measured per-step times."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402  (baseline leg only)

    logical, cores_per_socket, sockets = _host_cores()
    # the host CPU share that one GPU of the box comes with is 16 threads
    threads = min(16, os.cpu_count() or 1)
    mode = {"binomial": lbmx.CHECKPOINT_BINOMIAL, "dynamic": lbmx.CHECKPOINT_DYNAMIC,
            "static": lbmx.CHECKPOINT_STATIC}[mode_name]

    def measure(desc, grid, setup, steps_full, slots, init, kf, kg, threads_used):
        nx, ny, nz = grid
        N = nx * ny * nz
        o = oracle.Oracle(desc, threads=threads_used)
        setup(o, nx, ny, nz, steps_full)
        o.set_design(workloads.design_field(nx, ny, nz, seed=0))
        o.init_equilibrium(**init)
        t0 = time.perf_counter()
        o.forward(kf)
        tf = (time.perf_counter() - t0) / kf
        r = {"forward_mlups": N / tf / 1e6, "sample": f"{nx}x{ny}x{nz}: {kf} forward steps"}
        if slots is not None:
            t0 = time.perf_counter()
            o.gradient(kg, **init)
            tg = time.perf_counter() - t0
            ta = max(tg / kg - tf, 1e-12)
            rec, _ = lbmx.checkpoint_simulate(slots, steps_full, mode)
            t_pass = tf * (steps_full + rec) + ta * steps_full
            r.update({"reverse_step_mlups": N / ta / 1e6,
                      "reverse_phase_mlups": N * steps_full / (tf * rec + ta * steps_full) / 1e6,
                      "gradient_mlups_modelled": N * steps_full / t_pass / 1e6,
                      "recompute_sweeps": rec, "gradient_s_modelled": t_pass,
                      "sample": r["sample"] + f", a {kg}-step gradient with every state stored; the "
                                f"{steps_full}-step pass with {slots} slots ({mode_name}) modelled"})
        return r

    def d3(gx, gy, gz):
        return lbmx.make_desc(gx, gy, gz, lattice=lbmx.D3Q19, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN,
                              omega=1.6, interp_q=0.1, alpha_max=0.1)

    def duct(o, nx, ny, nz, steps):
        workloads.duct3d(o, nx, ny, nz, u_in=0.04, time_const=1.0 / steps)

    per = {}
    t_all = time.perf_counter()
    head = measure(d3(384, 192, 192), (384, 192, 192), duct, horizon_head, slots_head, {}, 3, 3, threads)
    per["duct_384x192x192_mrt_brinkman"] = head
    log(f"cpu baseline duct: forward {head['forward_mlups']:.2f} MLUPS, reverse steps {head['reverse_step_mlups']:.2f} "
        f"MLUPS, modelled gradient {head['gradient_mlups_modelled']:.2f} MLUPS on {threads} threads")
    sizes = {"channel_256x128_bgk_brinkman_forward": (200, None), "channel_1024x512_bgk_brinkman_modulated": (20, 10),
             "thermal_2048x1024_mrt": (6, 4)}
    for w in other_workloads(lbmx, workloads):
        nx, ny = w["grid"]
        kf, kg = sizes[w["name"]]
        r = measure(lbmx.make_desc(nx, ny, **w["desc"]), (nx, ny, 1),
                    lambda o, a, b, c, st, w=w: w["setup"](o, a, b, st), w["steps"], w["slots"], w["init"], kf,
                    kg or 1, threads)
        per[w["name"]] = r
        log(f"cpu baseline {w['name']}: {', '.join(f'{k} {v:.2f}' for k, v in r.items() if k.endswith('mlups') or k.endswith('modelled'))}")
    ref_adj = reference_adjoint_report(lbmx, workloads, oracle, threads)
    one = measure(d3(96, 48, 48), (96, 48, 48), duct, horizon_head, slots_head, {}, 4, 4, 1)
    log(f"cpu baseline duct on 1 thread (96x48x48): modelled gradient {one['gradient_mlups_modelled']:.3f} MLUPS")
    return {"value": head["gradient_mlups_modelled"], "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"D3Q19 MRT Brinkman duct 384x192x192, FP64: 3 forward steps and a 3-step gradient "
                      f"(every state stored) timed; the {horizon_head}-step pass with {slots_head} checkpoints "
                      f"({mode_name} placement, recomputation counted) modelled from those per-step times",
            "recompute": True, "modelled": True,
            "single_core": {"value": one["gradient_mlups_modelled"], "sample": "96x48x48 duct, same model, 1 thread"},
            "workloads": per, "reference_adjoint": ref_adj, "cpu": _cpu_model(), "host_logical_cpus": logical,
            "host_cores_per_socket": cores_per_socket, "host_sockets": sockets,
            "threads_note": "16 threads: the host CPU share that comes with one GPU of the box",
            "wall_s": round(time.perf_counter() - t_all, 1)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--horizon", type=int, default=3000)
    ap.add_argument("--slots", type=int, default=48)
    ap.add_argument("--placement", choices=["binomial", "dynamic", "static"], default="binomial",
                    help="checkpoint placement of every gradient pass")
    ap.add_argument("--no-placement-compare", action="store_true",
                    help="skip the extra gradient of the headline form with the other placement")
    ap.add_argument("--grid", type=str, default="384x192x192")
    ap.add_argument("--sweeps", type=int, default=50)
    ap.add_argument("--forms", type=str, default="fp64,fp16,fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=3000,
                    help="time steps of the batched robust-ensemble measurement (0: skip it)")
    ap.add_argument("--ensemble-slots", type=int, default=64)
    ap.add_argument("--ensemble-more-slots", type=int, default=96,
                    help="also time the batched pass with this many checkpoints (0: skip)")
    ap.add_argument("--no-design-pipeline", action="store_true")
    ap.add_argument("--no-mma", action="store_true")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="skip the 256x128, 1024x512 and 2048x1024 workloads")
    ap.add_argument("--only-other-workloads", type=str, default="",
                    help="comma-separated names: run only these of the other workloads and print their record "
                         "(for kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is designed but "
                                   "not built (slab decomposition along the slowest axis, one face exchange per step)", "n_gpus": args.gpus}))
        sys.exit(2)
    import lbmx
    import workloads
    placements = {"binomial": lbmx.CHECKPOINT_BINOMIAL, "dynamic": lbmx.CHECKPOINT_DYNAMIC,
                  "static": lbmx.CHECKPOINT_STATIC}
    if args.only_other_workloads:
        only = set(args.only_other_workloads.split(","))
        print(json.dumps(run_other_workloads(lbmx, workloads, placements[args.placement], only=only)), flush=True)
        return
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    forms = [f for f in args.forms.split(",") if f]
    mode = placements[args.placement]
    # the headline form also runs the reference's own (dynamic) placement once, or the binomial one
    cmp_name = "dynamic" if args.placement != "dynamic" else "binomial"
    head_form = "fp64" if "fp64" in forms else forms[0]
    res = {}
    for f in forms:
        res[f] = run_form(lbmx, workloads, f, nx, ny, nz, args.horizon, args.slots, args.steps, args.warmup,
                          args.sweeps, mode,
                          (cmp_name, placements[cmp_name]) if f == head_form and not args.no_placement_compare else None)
    # accuracy of the reduced-storage forms against the FP64 plain form on the same workload
    if "fp64" in res:
        g64, J64 = res["fp64"]["_sens"], res["fp64"]["objective"]
        for f in forms:
            if f == "fp64":
                continue
            g, J = res[f]["_sens"], res[f]["objective"]
            res[f]["vs_fp64"] = {
                "objective_rel_err": abs(J - J64) / abs(J64),
                "gradient_rel_err_max_norm": float(np.max(np.abs(g - g64)) / np.max(np.abs(g64))),
                "gradient_rel_err_l2": float(np.linalg.norm(g - g64) / np.linalg.norm(g64)),
                "mlups_ratio": res[f]["mlups_fwd_adj"] / res["fp64"]["mlups_fwd_adj"]}
            log(f"{f} vs fp64: {res[f]['vs_fp64']}")
    for f in forms:
        res[f].pop("_sens", None)
    head = res[head_form]
    traffic = read_traffic(f"{head_form}_fwd_traffic_bytes")
    roof = {"bound": "hbm", "achieved": round(head["fwd_gbs"], 1), "peak": PEAK_GBS, "unit": "GB/s",
            "frac": round(head["fwd_gbs"] / PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_fwd<D3Q19,MRT,%s> (forward sweep)" % head_form,
            "bytes_per_lu": head["fwd_bytes_per_lu"],
            "adjoint_kernel": {"achieved": round(head["adj_gbs"], 1), "frac": round(head["adj_gbs"] / PEAK_GBS, 4),
                               "bytes_per_lu": head["adj_bytes_per_lu"]},
            "gradient_pass": {"achieved": round(head["pass_gbs"], 1), "frac": round(head["pass_gbs"] / PEAK_GBS, 4)}}
    out = {
        "metric": METRIC,
        "value": round(head["mlups_fwd_adj"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_form == "fp64" else head_form,
        "data": "synthetic (SplitMix64 design field seed 0, 5x5x5 box filter)",
        "config": {"workload": f"D3Q19 MRT Brinkman duct {nx}x{ny}x{nz}, {args.horizon} time steps, "
                               f"forward+adjoint with {args.slots} HBM checkpoints ({args.placement} placement)",
                   "nodes": nx * ny * nz, "time_steps": args.horizon, "checkpoints": args.slots,
                   "placement": args.placement},
        "roofline": roof,
        "forms": {f: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for f, r in res.items()},
    }
    if "fp16" in res and "fp64" in res:
        out["fp16_over_fp64"] = round(res["fp16"]["mlups_fwd_adj"] / res["fp64"]["mlups_fwd_adj"], 3)
    if "fp32" in res and "fp64" in res:
        out["fp32_over_fp64"] = round(res["fp32"]["mlups_fwd_adj"] / res["fp64"]["mlups_fwd_adj"], 3)
    if args.ensemble_steps > 0:
        out["ensemble"] = run_ensemble(lbmx, workloads, args.ensemble_steps, args.ensemble_slots, min(args.warmup, 1),
                                       mode, args.placement, args.ensemble_more_slots)
    if not args.no_design_pipeline:
        out["design_pipeline"] = run_design_pipeline(lbmx, cpu=not args.no_cpu_baseline)
    if not args.no_mma:
        out["mma"] = run_mma(lbmx, cpu=not args.no_cpu_baseline)
    out["stream_copy_gbs"] = round(lbmx.stream_copy_gbs(), 1)
    log(f"stream copy ceiling {out['stream_copy_gbs']} GB/s")
    # the same bytes moved as 19 planes in and 19 planes out (the forward sweep's pattern)
    pc = lbmx.plane_copy_gbs(19, nx * ny * nz)
    out["roofline"]["plane_copy_gbs"] = round(pc, 1)
    out["roofline"]["frac_of_plane_copy"] = round(head["fwd_gbs"] / pc, 4)
    log(f"19-plane copy {pc:.1f} GB/s; forward sweep at {head['fwd_gbs'] / pc:.3f} of it")
    if "fp16" in out.get("forms", {}):
        # the 16-bit forward sweep's pattern: 19 planes of 2-byte values, two nodes per thread
        pc16 = lbmx.plane_copy16_gbs(19, nx * ny * nz)
        f16 = out["forms"]["fp16"]
        f16["plane_copy16_gbs"] = round(pc16, 1)
        f16["fwd_frac_of_plane_copy16"] = round(f16["fwd_gbs"] / pc16, 4)
        log(f"19-plane 16-bit copy {pc16:.1f} GB/s; FP16 forward sweep at {f16['fwd_gbs'] / pc16:.3f} of it")
    if not args.no_other_workloads:
        out["other_workloads"] = run_other_workloads(lbmx, workloads, mode)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(workloads, lbmx, args.horizon, args.slots, args.placement)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
