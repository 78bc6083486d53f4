"""GRAPE fidelity+gradient evaluations per second on one MI355X.

Workload: 4-qubit (d=16) closed-system gate synthesis with a 256-member
robustness ensemble (16 detunings x 16 control-amplitude scales), 8 x/y
drives, 2000 slices of 1 us, Haar-random 16x16 target; synthetic seeded
inputs. One step = one ensemble F and its gradient dF/du [8 x 2000] at one
control point, i.e. one ensemble_value_and_grad of build_ensemble_from_systems
(src/Optimization/Ensemble/EnsembleObjective.jl:272-307), with the controls
already resident in HBM (pr_eval_device on the plan's buffers).

Printed JSON line: the metric, the roofline of the dominant kernel (kernel
time from HIP events recorded by the library on its launch stream), the
accuracy of the device result against the FP64 CPU oracle on a member sample,
and the oracle timed on this host's cores (the CPU baseline). The FP64 run
(the default) also times the mixed-precision form on the same inputs and
reports it, with its own accuracy, under "mixed_precision_form".

    python bench.py [--gpus 1] [--steps 30] [--warmup 3] [--precision fp64|mixed] [--fp64-only]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pulsar.jl_amd"))

import numpy as np  # noqa: E402

from pulsar_rocm import native, workloads  # noqa: E402
from pulsar_rocm.problem import PR_FP64, PR_MIXED, PR_SPLIT16  # noqa: E402

# published dense matrix-core peaks of the MI355X (TFLOP/s)
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0}

# precision forms of the headline workload: (pr_options.precision, dtype)
FORMS = {"fp64": (PR_FP64, "f64"), "mixed_fp32": (PR_MIXED, "f64+f32"), "split16": (PR_SPLIT16, "f64+f16")}
WORK_MODEL = {
    "fp64": "complex products on FP64 matrix cores, 6 d^3 flops each (three real products)",
    "mixed_fp32": "exp and walk products on FP32 matrix cores and the chain product on FP64, 6 d^3 flops each",
    "split16": "FP16x2 split products on FP16 matrix cores, 8 d^3 flops per part-product (13 per slice in "
               "k_segment, 6 in k_gradient), no FP64 chain product; 2 FP64 products per work item (projection)",
}
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the two main kernels on this workload, measured with
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by what the same counters
# read for a 256 MiB device-to-device copy (FETCH x 2.0, WRITE x 1.0); the
# commands and per-kernel figures are in profiles/grape_four_qubit_ensemble_hbm_traffic.md.
# FP64: k_segment writes the U_k^T images (256 x 2000 x 4 KiB = 2.10 GB) and
# k_gradient reads them back once; split16: the split images of D_k (2 KiB,
# 1.05 GB), profiles/grape_four_qubit_ensemble_split16_hbm_traffic.md.
# k_track_walk (four_qubit_tracking) reads the FP64 images once too
# (profiles/grape_four_qubit_tracking_hbm_traffic.md).
MEASURED_TRAFFIC_BYTES = {("fp64", "k_segment"): 0.022e9 + 2.131e9, ("fp64", "k_gradient"): 2.141e9 + 0.037e9,
                          ("split16", "k_segment"): 0.010e9 + 1.083e9, ("split16", "k_gradient"): 1.092e9 + 0.033e9,
                          "k_track_walk": 2.166e9 + 0.040e9}


def mfma_work(kernel, form, counts, units, d, items):
    """Matrix-core flops one launch of the dominant kernels needs on this
    workload, per core type, from the exp counts the library reports
    (pr_plan_exp_counts: member-slices at degree 6, at degree 9, squarings,
    member-slices redone in FP64).

    fp64: every product on FP64 matrix cores, 6 d^3 each (three real products,
    Gauss). k_segment runs the exp (Paterson-Stockmeyer: 3 products at degree 6,
    4 at degree 9, one per squaring) and the chain product U_k T per
    member-slice; k_gradient the walk's M_k = C_k U_k^dag and C_{k+1} = U_k M_k.
    mixed_fp32: the degree-6 exp and both walk products on FP32 matrix cores
    (6 d^3 each), the chain product on FP64.
    split16: the exp's products X X (four FP16 part-products), X X^2 and X^3 Q
    (three each), the segment product's E_k F (three, every slice but the
    first of a work item) and the walk's two products (three each), as
    16x16x32 FP16 matrix products of 8 d^3 flops per part-product (the real 2x2
    block form); no FP64 chain product (segments of at most 128 slices), and
    two FP64 products per work item for the unitary projection of its segment
    product.
    Member-slices redone in FP64 count their FP64 exp and walk products."""
    n6, n9, nsq, nredo = (int(c) for c in counts)
    g = 6 * d ** 3
    if form == "fp64":
        if kernel == "k_segment":
            return {"fp64": g * (3 * n6 + 4 * n9 + nsq + units)}
        return {"fp64": g * 2 * units}
    low = units - nredo  # member-slices that stayed in the reduced-precision form
    redo_exp = 3 * n6 + 4 * n9 + nsq - 3 * low
    if form == "mixed_fp32":
        if kernel == "k_segment":
            return {"fp32": g * 3 * low, "fp64": g * (redo_exp + units)}
        return {"fp32": g * 2 * low, "fp64": g * 2 * nredo}
    if kernel == "k_segment":
        # (+ the unitary projection of each work item's segment product: 2 FP64 products)
        return {"fp16": 8 * d ** 3 * ((4 + 3 + 3) * low + 3 * max(low - items, 0)),
                "fp64": g * (redo_exp + nredo + 2 * items)}
    return {"fp16": 8 * d ** 3 * 3 * 2 * low, "fp64": g * 2 * nredo}


def roofline_of(work, ms):
    """achieved TFLOP/s, the peak of this mix of core types (the flops over the
    time they need at each core's peak) and their ratio, for a launch of ms."""
    flops = sum(work.values())
    t_peak = sum(v / (MFMA_PEAK_TFLOPS[k] * 1e12) for k, v in work.items())
    achieved = flops / (ms * 1e-3) / 1e12
    peak = flops / t_peak / 1e12
    return achieved, peak, t_peak / (ms * 1e-3)


def vector_flops(d, nc):
    """VALU flops per member-slice outside the products: assembly of the
    generator (n_c real-weighted complex axpys, 4 n_c d^2) in k_segment and the
    n_c control traces (8 n_c d^2) in k_gradient."""
    return {"k_segment": 4 * nc * d ** 2, "k_gradient": 8 * nc * d ** 2}


def survey_flops(d, nc):
    """The fixed comparison model of the workload description: 72 d^3 + 16 n_c d^2
    per member-slice (nine 8 d^3 complex products, six of them for exp)."""
    return 72 * d ** 3 + 16 * nc * d ** 2


def cpu_baseline(p, u, n_sample, threads):
    """The FP64 oracle (reference-faithful: separate fidelity and gradient
    passes, as the per-member closures of EnsembleObjective do) on a member
    sample; returns (evals/s scaled to the full ensemble, seconds, F, G)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    q, uu = workloads.shrink(p, u, n_members=n_sample)
    t0 = time.perf_counter()
    F, G = oracle.evaluate(q, uu, separate_fg=True, n_threads=threads)
    dt = time.perf_counter() - t0
    return (n_sample / p.n_members) / dt, dt, F, G, q


def cpu_forms(p, u, threads_arg, threads):
    """The oracle's other timed forms: one fused F and G pass on all threads
    (the whole ensemble), and the reference-faithful separate passes on one
    thread (8 members); rates scaled to the full ensemble, with the host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    out = {}
    t0 = time.perf_counter()
    oracle.evaluate(p, u, separate_fg=False, n_threads=threads_arg)
    dt = time.perf_counter() - t0
    out["fused_fg_all_threads"] = {"value": round(1.0 / dt, 5), "unit": "evals/s", "cores": threads,
                                   "sample": f"all {p.n_members} members, {dt:.2f} s"}
    n = 8
    q, uu = workloads.shrink(p, u, n_members=n)
    t0 = time.perf_counter()
    oracle.evaluate(q, uu, separate_fg=True, n_threads=1)
    dt = time.perf_counter() - t0
    out["separate_fg_one_thread"] = {"value": round((n / p.n_members) / dt, 6), "unit": "evals/s", "cores": 1,
                                     "sample": f"{n} of {p.n_members} members, {dt:.2f} s, scaled"}
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    out["host"] = {"cpu_model": model, "nproc": os.cpu_count(), "threads_used": threads}
    return out


def calibrate_copy(mib):
    """A plain device-to-device copy of known size through the HIP runtime, so
    that FETCH_SIZE / WRITE_SIZE of the copy kernel calibrate the counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": n, "copies": 3}))


def other_workload(args):
    """Timing of the other GRAPE configurations (parity is covered by the tests):
    evaluations per second of the whole batch, device-resident controls."""
    if args.workload == "three_qubit_lindblad":
        p, u = workloads.three_qubit_lindblad()
        probs, us = [p], [u]
        desc = "three_qubit_lindblad: N=8 (64x64 Liouvillian), 64 members, 6 controls, 1000 slices, dt=2ns, 64 pairs"
        flops = 8 * 64 ** 3 * (6 + 2 + p.n_ctrl) * p.n_members * p.n_t
    elif args.workload == "uneven_batch":
        batch = workloads.uneven_batch()
        probs, us = [b[0] for b in batch], [b[1] for b in batch]
        desc = "uneven_batch: 1024 gate problems, d in {2..64} with p ~ 1/d, 200..2000 slices, 2 controls per qubit"
        flops = sum((72 * q.dim ** 3 + 16 * q.n_ctrl * q.dim ** 2) * q.n_t * q.n_members for q in probs)
    elif args.workload == "spin_chain_128":
        p, u = workloads.spin_chain_128()
        probs, us = [p], [u]
        desc = "spin_chain_128: d=128 Heisenberg chain, 64 members, 14 controls, 500 slices, gate target"
        flops = (72 * 128 ** 3 + 16 * p.n_ctrl * 128 ** 2) * p.n_members * p.n_t
    elif args.workload == "four_qubit_tracking":
        p, u = workloads.four_qubit_tracking()
        probs, us = [p], [u]
        desc = (f"four_qubit_tracking: d=16, 256 members, 8 controls, 2000 slices, {p.n_pairs} initial states, "
                f"{p.n_track} density-matrix checkpoints")
        # exps + segment chain (5 complex products) and the walk (3 per slice)
        flops = (64 * 16 ** 3 + 16 * p.n_ctrl * 16 ** 2) * p.n_members * p.n_t
    else:
        p, u = workloads.single_qubit_state()
        probs, us = [p], [u]
        desc = "single_qubit_state: d=2, 100 slices, state transfer"
        flops = (72 * 8 + 16 * 2 * 4) * 100
    native.init(0)
    prec = {"fp64": PR_FP64, "mixed": PR_MIXED, "split16": PR_SPLIT16}[args.precision]
    plan = native.Plan(probs, precision=prec, segment=args.segment)
    plan.upload_controls(us)
    d_u, d_F, d_G = plan.buffers()
    for _ in range(args.warmup):
        plan.eval_device(d_u, d_F, d_G)
    plan.synchronize()
    # kernel breakdown with per-launch events (direct launches) ...
    plan.set_timing(True)
    plan.kernel_times()
    for _ in range(args.steps):
        plan.eval_device(d_u, d_F, d_G)
    plan.synchronize()
    kt = plan.kernel_times()
    plan.set_timing(False)
    counts = plan.exp_counts()
    # ... and the throughput of replayed evaluation graphs
    plan.eval_device(d_u, d_F, d_G)
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.eval_device(d_u, d_F, d_G)
    plan.synchronize()
    elapsed = time.perf_counter() - t0
    # the host-buffer entry point pr_eval (one packed copy of the controls up,
    # one of the gradients down, through a pinned staging buffer), called as
    # a ccall would, with the per-problem host arrays made once ...
    abi_ms = c_abi_eval_ms(plan, probs, us, args.steps)
    # ... and through the Python mirror, which also converts every problem's
    # controls and gradient between numpy and column-major buffers per call
    plan.eval(us)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.eval(us)
    host_elapsed = time.perf_counter() - t0
    plan.close()
    rate = args.steps / elapsed
    line = {"workload": args.workload, "description": desc, "evals_per_s": round(rate, 3),
            "ms_per_eval": round(1e3 * elapsed / args.steps, 4), "steps": args.steps,
            "algorithmic_tflops": round(flops * rate / 1e12, 3),
            "kernel_ms_per_eval": {k: round(v[0] / args.steps, 4) for k, v in kt.items()},
            "pr_eval_host_buffers_ms_per_eval": round(abi_ms, 4),
            "python_mirror_eval_ms_per_eval": round(1e3 * host_elapsed / args.steps, 4)}
    if args.workload == "uneven_batch":
        # the same batch without packing small problems (pr_options.pack = 0);
        # its per-problem exp counts are the work model's (a pack reports its
        # exps once, for all of its problems)
        line["unpacked"], counts = unpacked_timing(probs, us, args, prec)
    if prec == PR_FP64 and args.workload != "four_qubit_tracking":
        line["roofline"] = fp64_roofline(probs, counts, kt, args.steps)
    if args.workload in ("spin_chain_128", "three_qubit_lindblad", "four_qubit_tracking"):
        line["recompute_mode"] = recompute_timing(probs, us, args)
    if args.workload == "uneven_batch":
        line["autotuned"] = autotuned_timing(probs, us, args, prec)
    if args.cpu_sample > 0 and args.workload != "four_qubit_tracking":
        line["accuracy"] = sample_accuracy(args.workload, probs, us, prec)
    if args.workload == "four_qubit_tracking" and "gradient" in kt:
        # the per-segment backward walk reads every U_k image once (16 x 16
        # complex FP64, 4 KiB) and writes the member's gradient row (n_c
        # doubles) per member-slice: HBM is its bound (its MFMA work, 8 d^2 2P
        # + 8 d^2 P + 8 per operator nonzero flops, is 13.1 GFLOP, 0.17 ms at peak)
        walk_bytes = (16 * 16 * 16 + 8 * p.n_ctrl) * p.n_members * p.n_t
        walk_s = kt["gradient"][0] / args.steps * 1e-3
        line["roofline"] = {"bound": "hbm", "kernel": "k_track_walk", "achieved": round(walk_bytes / walk_s / 1e9, 1),
                            "peak": 8000.0, "unit": "GB/s", "frac": round(walk_bytes / walk_s / 8e12, 4),
                            "traffic": MEASURED_TRAFFIC_BYTES["k_track_walk"],
                            "algorithmic_bytes_per_launch": walk_bytes}
    if args.workload == "four_qubit_tracking" and args.cpu_sample > 0:
        # the tracking oracle (density matrices, one thread) on a member sample
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        n = min(args.cpu_sample, 16)  # about 0.65 s of one core per member
        q, uu = workloads.shrink(p, u, n_members=n)
        t0 = time.perf_counter()
        Fo, Go = O.tracking_evaluate(q, uu)
        cpu_s = time.perf_counter() - t0
        qplan = native.Plan([q], precision=prec)
        Fd, Gd = qplan.eval([uu])
        qplan.close()
        line["accuracy"] = {"sample": f"{n} members, all {p.n_t} slices, against the tracking oracle",
                            "abs_dF": float(abs(Fd[0] - Fo)),
                            "max_rel_dG": float(np.abs(Gd[0] - Go).max() / np.abs(Go).max())}
        line["cpu_baseline"] = {"value": round(n / (cpu_s * p.n_members), 6), "unit": "evals/s", "cores": 1,
                                "kind": "port", "sample": f"{n} of {p.n_members} members, all {p.n_t} slices, "
                                f"{cpu_s:.1f} s; rate scaled to the full ensemble"}
    print(json.dumps(line))


def unpacked_timing(probs, us, args, prec):
    """The batch with every problem its own work (pr_options.pack = 0):
    evaluations/s, per-kernel times, and the per-problem exp counts."""
    plan = native.Plan(probs, precision=prec, segment=args.segment, pack=0)
    try:
        plan.upload_controls(us)
        d_u, d_F, d_G = plan.buffers()
        for _ in range(args.warmup):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        plan.set_timing(True)
        plan.kernel_times()
        for _ in range(args.steps):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        kt = plan.kernel_times()
        plan.set_timing(False)
        counts = plan.exp_counts()
        plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        el = time.perf_counter() - t0
    finally:
        plan.close()
    return ({"evals_per_s": round(args.steps / el, 3), "ms_per_eval": round(1e3 * el / args.steps, 4),
             "kernel_ms_per_eval": {k: round(v[0] / args.steps, 4) for k, v in kt.items()}}, counts)


def autotuned_timing(probs, us, args, prec):
    """The same batch with the library's segment-length autotuner
    (PR_SEGMENT_AUTOTUNE): plan-creation time, evaluations/s and the segment
    count it chose per dimension."""
    from pulsar_rocm.problem import PR_SEGMENT_AUTOTUNE
    t0 = time.perf_counter()
    plan = native.Plan(probs, precision=prec, segment=PR_SEGMENT_AUTOTUNE)
    create_s = time.perf_counter() - t0
    try:
        plan.upload_controls(us)
        d_u, d_F, d_G = plan.buffers()
        for _ in range(args.warmup):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        el = time.perf_counter() - t0
        S = plan.shapes()
    finally:
        plan.close()
    seglen, regs = {}, {}
    for q, sh in zip(probs, S):
        seglen.setdefault(q.dim, set()).add(int(sh[0]))
        regs.setdefault(q.dim, set()).add(int(sh[2]))
    return {"evals_per_s": round(args.steps / el, 3), "ms_per_eval": round(1e3 * el / args.steps, 4),
            "plan_create_s": round(create_s, 3),
            "slices_per_segment_by_dim": {int(k): sorted(v)[:3] for k, v in sorted(seglen.items())},
            "trace_operand_in_registers_by_dim": {int(k): sorted(v) for k, v in sorted(regs.items())}}


def recompute_timing(probs, us, args, prec=PR_FP64):
    """The same evaluation without stored U_k images (store_props = 0): the
    gradient walk recomputes each U_k from segment checkpoints. Lindblad
    problems also keep their forward states only for the work items of one
    launch window (segments of at most 32 slices, 8192 / B^2 work items), in
    place of every slice's."""
    plan = native.Plan(probs, store_props=0, segment=args.segment, precision=prec)
    try:
        plan.upload_controls(us)
        d_u, d_F, d_G = plan.buffers()
        for _ in range(args.warmup):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        plan.set_timing(True)
        plan.kernel_times()
        for _ in range(args.steps):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        kt = plan.kernel_times()
        plan.set_timing(False)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.eval_device(d_u, d_F, d_G)
        plan.synchronize()
        el = time.perf_counter() - t0
        D = plan.eval(us, diag=True)[2]
    finally:
        plan.close()
    # FP64 images: 4 KiB per 16 x 16 block; split-form images: 2 KiB (d <= 16)
    per_block = 2048 if prec == PR_SPLIT16 else 4096
    images = sum(q.n_members * q.n_t * ((q.dim + 15) // 16) ** 2 * per_block for q in probs)
    out = {"evals_per_s": round(args.steps / el, 3), "ms_per_eval": round(1e3 * el / args.steps, 3),
           "kernel_ms_per_eval": {k: round(v[0] / args.steps, 4) for k, v in kt.items()},
           "segments_per_member": int(D[0, 3]), "image_bytes_saved": int(images)}
    from pulsar_rocm.problem import PR_LINDBLAD
    if any(q.dynamics == PR_LINDBLAD for q in probs):
        saved = 0
        for q, dd in zip(probs, D):
            if q.dynamics != PR_LINDBLAD:
                continue
            nb = ((q.dim + 15) // 16) ** 2
            S = int(dd[3])
            L = -(-q.n_t // S)
            slots = min(q.n_members * S, max(1, 8192 // nb))
            saved += q.n_members * q.n_t * nb * 4096 - slots * 2 * L * nb * 4096
        out["forward_state_bytes_saved"] = int(saved)
    return out


def c_abi_eval_ms(plan, probs, us, steps):
    """ms per pr_eval call with host buffers, the per-problem column-major
    control and gradient arrays and their pointer tables made beforehand."""
    import ctypes
    from pulsar_rocm.problem import controls_cm
    L = native.load()
    n = len(probs)
    dp = ctypes.POINTER(ctypes.c_double)
    bufs = [controls_cm(u) for u in us]
    gbufs = [np.zeros(q.n_ctrl * q.n_t) for q in probs]
    cptrs = (dp * n)(*[b.ctypes.data_as(dp) for b in bufs])
    gptrs = (dp * n)(*[g.ctypes.data_as(dp) for g in gbufs])
    F = np.zeros(n)
    native.check(L.pr_eval(plan.handle, cptrs, F.ctypes.data_as(dp), gptrs, None))
    t0 = time.perf_counter()
    for _ in range(steps):
        native.check(L.pr_eval(plan.handle, cptrs, F.ctypes.data_as(dp), gptrs, None))
    return 1e3 * (time.perf_counter() - t0) / steps


def fp64_roofline(probs, counts, kt, steps):
    """FP64 matrix-core work of the segment and gradient kernels per
    evaluation, from the exp counts: each complex product of dimension D (the
    generator's: d closed, d^2 Lindblad) costs 6 D^3 flops as three real
    products (Gauss, D <= 64) or 8 D^3 as four (D > 64). k_segment: the exp's
    products (3 at degree 6, 4 at degree 9, one per squaring) and the chain
    product per member-slice; k_gradient: 2 products per member-slice for
    unitary dynamics (M = C U^dag, C' = U M), 3 for Lindblad (forward state,
    P Lambda^dag, U^dag Lambda). Traces and assembly (VALU) are not counted.
    The dominant kernel class is the one with more time per evaluation."""
    from pulsar_rocm.problem import PR_LINDBLAD as LIND
    seg = grad = 0.0
    for q, c in zip(probs, counts):
        D = q.dim
        g = (6 if D <= 64 else 8) * D ** 3
        units = q.n_members * q.n_t
        seg += g * (3 * c[0] + 4 * c[1] + c[2] + units)
        grad += g * (3 if q.dynamics == LIND else 2) * units
    seg_ms, grad_ms = kt["segment"][0] / steps, kt["gradient"][0] / steps
    kern, flops, ms = ("k_segment", seg, seg_ms) if seg_ms >= grad_ms else ("k_gradient", grad, grad_ms)
    achieved = flops / (ms * 1e-3) / 1e12
    return {"bound": "mfma", "kernel": kern, "achieved": round(achieved, 3), "peak": MFMA_PEAK_TFLOPS["fp64"],
            "unit": "TFLOP/s", "frac": round(achieved / MFMA_PEAK_TFLOPS["fp64"], 4), "traffic": None,
            "kernel_ms": round(ms, 4), "flops_per_eval": int(flops),
            "work_model": "FP64 complex products, 6 D^3 flops (Gauss, D <= 64) or 8 D^3 (D > 64) each; "
                          "k_segment exp + chain per member-slice, k_gradient 2 (unitary) or 3 (Lindblad); "
                          "all launches of the kernel class in one evaluation"}


def sample_accuracy(workload, probs, us, prec):
    """|dF|, max relative dG and the unitarity defect of the device on a
    bounded sample of the workload against the FP64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if workload == "uneven_batch":
        # the first 2 problems of each dimension, taken from an evaluation of
        # the whole batch (the benchmarked plan, small problems packed)
        pick = []
        for d in sorted({q.dim for q in probs}):
            pick += [i for i, q in enumerate(probs) if q.dim == d][:2]
        plan = native.Plan(probs, precision=prec)
        Fa, Ga, Da = plan.eval(us, diag=True)
        plan.close()
        sp, su = [probs[i] for i in pick], [us[i] for i in pick]
        Fd, Gd, D = Fa[pick], [Ga[i] for i in pick], Da[pick]
        desc = (f"{len(sp)} problems (the first 2 of each dimension) of the whole-batch evaluation, "
                "full length")
    else:
        if workload in ("three_qubit_lindblad", "spin_chain_128"):
            q, uu = workloads.shrink(probs[0], us[0], n_members=2)
            sp, su = [q], [uu]
            desc = f"2 of {probs[0].n_members} members, all {q.n_t} slices"
        else:
            sp, su = probs, us
            desc = "the whole workload"
        plan = native.Plan(sp, precision=prec)
        Fd, Gd, D = plan.eval(su, diag=True)
        plan.close()
    dF = dG = 0.0
    for i, (q, uu) in enumerate(zip(sp, su)):
        Fo, Go = O.evaluate(q, uu)
        dF = max(dF, abs(Fd[i] - Fo))
        dG = max(dG, float(np.abs(Gd[i] - Go).max() / np.abs(Go).max()))
    from pulsar_rocm.problem import PR_CLOSED as CLOSED
    closed = [i for i, q in enumerate(sp) if q.dynamics == CLOSED]
    return {"sample": desc, "abs_dF": float(dF), "max_rel_dG": dG,
            "unitarity_defect": float(np.max(D[closed, 0])) if closed else None}


def mr_workload(args):
    """MR state transfer of one spin over 64 offsets x 16 RF power levels
    (1024 members, 500 slices of 2 us): the grape_state_kernel evaluation
    with the members packed 8 per 16 x 16 block by the Python mirror
    (mr.PackedMembers) and by the library (pr_options.pack), and unpacked."""
    from pulsar_rocm import mr
    ctrl, w = workloads.mr_spin_ensemble()
    native.init(0)
    out = {"workload": "mr_spin_ensemble",
           "description": "1 spin, 64 offsets x 16 power levels (1024 members), 2 controls, 500 slices, 1 pair"}
    for pack in (True, False):
        kp = mr.KernelPlan(ctrl, pack=pack)
        try:
            for _ in range(args.warmup):
                kp(w)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                kp(w)
            dt = (time.perf_counter() - t0) / args.steps
        finally:
            kp.close()
        out["packed" if pack else "unpacked"] = {"evals_per_s": round(1 / dt, 2), "ms_per_eval": round(1e3 * dt, 4)}
    # the same ensemble handed to the library as it is, which packs its
    # members itself (pr_options.pack)
    plan = native.Plan([ctrl.problem()])
    try:
        for _ in range(args.warmup):
            plan.eval([w])
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.eval([w])
        dt = (time.perf_counter() - t0) / args.steps
    finally:
        plan.close()
    out["packed_by_the_library"] = {"evals_per_s": round(1 / dt, 2), "ms_per_eval": round(1e3 * dt, 4)}
    out["note"] = "host-buffer evaluations (controls up, gradient down each call)"
    print(json.dumps(out))


def accuracy(q, u, Fo, Go, prec, members):
    """The form's F, G and unitarity defect on the oracle's member sample q
    against the oracle's Fo, Go."""
    qplan = native.Plan([q], precision=prec)
    Fd, Gd, D = qplan.eval([u], diag=True)
    qplan.close()
    return {"members": members, "abs_dF": float(abs(Fd[0] - Fo)),
            "max_rel_dG": float(np.abs(Gd[0] - Go).max() / np.abs(Go).max()),
            "max_abs_dG": float(np.abs(Gd[0] - Go).max()), "max_abs_G": float(np.abs(Go).max()),
            "unitarity_defect": float(D[0, 0]), "F": float(Fd[0])}


def lbfgs_check(p, u_start, prec):
    """84 device L-BFGS iterations on p from u_start with the form's F and G;
    returns the FP64 fidelity of the result."""
    cfg = native.LbfgsConfig(max_iter=84, convergence_tol=1e-15, step_size=0.05 * np.abs(u_start).max(),
                             gradient_norm_tol=1e-15, lbfgs_memory=10, amplitude_limit=0.0)
    plan = native.Plan([p], precision=prec)
    t0 = time.perf_counter()
    best, fh, _, res = native.lbfgs_optimize_device(plan, u_start, cfg)
    el = time.perf_counter() - t0
    plan.close()
    ref = native.Plan([p], precision=PR_FP64)
    F64 = ref.eval([best], grad=False)[0][0]
    ref.close()
    return {"iterations": int(res.n_iterations), "seconds": round(el, 3), "F_in_fp64": float(F64),
            "F_own": float(fh[-1])}


def timed_form(p, u, prec, args):
    """Warm-up, then K evaluations with the library's per-kernel events and K
    without them (the headline time is the faster loop). Returns the open plan,
    the elapsed seconds of K evaluations and the per-kernel times."""
    plan = native.Plan([p], precision=prec, segment=args.segment)
    plan.upload_controls([u])
    d_u, d_F, d_G = plan.buffers()
    for _ in range(args.warmup):
        plan.eval_device(d_u, d_F, d_G)
    plan.synchronize()

    plan.set_timing(True)
    plan.kernel_times()  # reset
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.eval_device(d_u, d_F, d_G)
    plan.synchronize()
    elapsed = time.perf_counter() - t0
    kt = plan.kernel_times()
    plan.set_timing(False)
    counts = plan.exp_counts()

    # the same timed loop without event records, for the headline number
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.eval_device(d_u, d_F, d_G)
    plan.synchronize()
    elapsed_plain = time.perf_counter() - t0
    return plan, min(elapsed, elapsed_plain), kt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "split16"],
                    help="precision of the --workload timing lines")
    ap.add_argument("--forms", default="all",
                    help="comma-separated precision forms of the headline workload (fp64, mixed_fp32, split16)")
    ap.add_argument("--workload", default="four_qubit_ensemble",
                    choices=["four_qubit_ensemble", "three_qubit_lindblad", "uneven_batch", "single_qubit_state",
                             "spin_chain_128", "four_qubit_tracking", "mr_spin_ensemble"],
                    help="the headline metric is four_qubit_ensemble; the others print timing lines of their own")
    ap.add_argument("--cpu-sample", type=int, default=256,
                    help="members timed for the CPU baseline (0: skip); the default is the whole ensemble")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--segment", type=int, default=0,
                    help="slices per time segment of the chain scan (0: the library's choice)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.workload == "mr_spin_ensemble":
        mr_workload(args)
        return
    if args.workload != "four_qubit_ensemble":
        other_workload(args)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print(f"bench.py: only the single-GPU measurement exists (asked for {max(args.gpus, world)} GPUs); "
                  "the ensemble would shard members across GPUs with one all-reduce of (F, G) per step, "
                  "which this build does not run.", file=sys.stderr)
        sys.exit(2)

    p, u = workloads.four_qubit_ensemble()
    native.init(0)
    M, nt, d, nc = p.n_members, p.n_t, p.dim, p.n_ctrl
    units = M * nt
    # the oracle (reference algorithm, FP64) on the same inputs: the accuracy
    # reference of every form, and the CPU baseline
    cpu = None
    if args.cpu_sample > 0:
        cpu_rate, cpu_s, Fo, Go, q = cpu_baseline(p, u, args.cpu_sample, args.cpu_threads)
        cpu = (cpu_rate, cpu_s, Fo, Go, q)
    # the GRAPE-optimized point (tests/golden/c2_optimized.npz: the nominal
    # system's propagator at the reference controls as the target, controls
    # from the device's ensemble L-BFGS, ensemble F = 0.99967): accuracy near
    # F = 1 and the time there
    opt = None
    if args.cpu_sample > 0:
        fx = np.load(os.path.join(ROOT, "tests", "golden", "c2_optimized.npz"))
        po, uo = workloads.shrink(p, u)
        po.U_target = fx["target"]
        uo = fx["controls"]
        qo, _ = workloads.shrink(po, uo, n_members=args.cpu_sample)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        Foo, Goo = oracle.evaluate(qo, uo, n_threads=args.cpu_threads)
        opt = (po, uo, Foo, Goo, qo)
    forms = {}
    names = [n for n in FORMS if args.forms == "all" or n in args.forms.split(",")]
    for name in names:
        prec = FORMS[name][0]
        plan, elapsed, kt = timed_form(p, u, prec, args)
        counts = plan.exp_counts()[0]
        D0 = plan.eval([u], diag=True)[2][0]  # diag[3]: time segments per member
        plan.close()
        seg_ms = kt["segment"][0] / max(kt["segment"][1], 1)
        grad_ms = kt["gradient"][0] / max(kt["gradient"][1], 1)
        dominant, dom_ms = ("k_segment", seg_ms) if seg_ms >= grad_ms else ("k_gradient", grad_ms)
        work = mfma_work(dominant, name, counts, units, d, M * int(D0[3]))
        achieved, peak, frac = roofline_of(work, dom_ms)
        f = {"value": round(args.steps / elapsed, 3), "ms_per_step": round(1e3 * elapsed / args.steps, 4),
             "dtype": FORMS[name][1],
             "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items()},
             "segments_per_member": int(D0[3]),
             "exp_counts": {"deg6": int(counts[0]), "deg9": int(counts[1]), "squarings": int(counts[2]),
                            "fp64_redone": int(counts[3])},
             "roofline": {"bound": "mfma", "kernel": dominant, "achieved": round(achieved, 3),
                          "peak": round(peak, 1), "unit": "TFLOP/s", "frac": round(frac, 4),
                          "traffic": MEASURED_TRAFFIC_BYTES.get((name, dominant)),
                          "kernel_ms": round(dom_ms, 4),
                          "flops_per_launch": {k: int(v) for k, v in work.items()},
                          "work_model": WORK_MODEL[name] + f"; {units} member-slices per launch, exp counts "
                                        f"{[int(c) for c in counts]}; VALU flops outside the products "
                                        f"({vector_flops(d, nc)[dominant]} per member-slice) not counted"}}
        if cpu is not None:
            _, _, Fo, Go, q = cpu
            f["accuracy"] = accuracy(q, u, Fo, Go, prec, args.cpu_sample)
            f["meets_gate"] = f["accuracy"]["abs_dF"] <= 1e-9 and f["accuracy"]["max_rel_dG"] <= 1e-6
        if opt is not None:
            po, uo, Foo, Goo, qo = opt
            plan, el_o, kt_o = timed_form(po, uo, prec, args)
            plan.close()
            f["at_optimized_point"] = {"value": round(args.steps / el_o, 3),
                                       "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt_o.items()},
                                       "accuracy": accuracy(qo, uo, Foo, Goo, prec, args.cpu_sample)}
            # at the ensemble optimum the members' gradients cancel in the mean
            # (max|G| there is ~1e-4 of its value at the benchmark point), so
            # the gradient's relative error is gated at the benchmark point
            # and only reported here; F is gated at both
            a = f["at_optimized_point"]["accuracy"]
            f["meets_gate"] = f.get("meets_gate", True) and a["abs_dF"] <= 1e-9
            # what the gradient is for: 84 L-BFGS iterations from the
            # reference controls with this form's F and G; F of the result in FP64
            f["at_optimized_point"]["lbfgs_84"] = lbfgs_check(po, u, prec)
        if name in ("fp64", "split16"):
            # the same evaluation without stored images (store_props = 0)
            f["recompute_mode"] = recompute_timing([p], [u], args, prec)
        forms[name] = f
    # the headline: the fastest form within the gate against the oracle: |dF|
    # <= 1e-9 at the benchmark point and at the optimized point, max relative
    # gradient error <= 1e-6 at the benchmark point; FP64 without an oracle run
    gated = [n for n in names if forms[n].get("meets_gate", n == "fp64")]
    head = max(gated, key=lambda n: forms[n]["value"]) if gated else names[0]
    h = forms[head]
    out = {
        "metric": "GRAPE fidelity+gradient evals/s (4q x 256 ens) at |dF|<=1e-9 vs FP64",
        "value": h["value"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": h["dtype"],
        "data": "synthetic (seeded numpy: random controls, Haar target)",
        "config": {"workload": "four_qubit_ensemble: d=16, 256 members (16 detuning x 16 amplitude), "
                               "8 controls, 2000 slices, dt=1us, gate target",
                   "members": M, "slices": nt, "dim": d, "controls": nc, "form": head},
        "roofline": h["roofline"],
        "survey_model_tflops": round(survey_flops(d, nc) * units * h["value"] / 1e12, 3),
        "kernel_ms_per_step": h["kernel_ms_per_step"],
        "gate": "vs the FP64 oracle: |dF| <= 1e-9 at these controls and at GRAPE-optimized controls "
                "(tests/golden/c2_optimized.npz, F = 0.99967); max|dG|/max|G| <= 1e-6 at these controls",
    }
    if cpu is not None:
        cpu_rate, cpu_s = cpu[0], cpu[1]
        threads = args.cpu_threads or os.cpu_count()
        out["cpu_baseline"] = {"value": round(cpu_rate, 5), "unit": "evals/s", "cores": threads, "kind": "port",
                               "sample": f"{args.cpu_sample} of {M} members, all {nt} slices, FP64 oracle with "
                                         f"separate F and G passes, {threads} OpenMP threads, {cpu_s:.2f} s"}
        out["accuracy"] = h["accuracy"]
        out["cpu_forms"] = cpu_forms(p, u, args.cpu_threads, threads)
    out["propagator_steps_per_s"] = round(h["value"] * M * nt, 1)
    if "fp64" in forms:
        for n in names:
            forms[n]["vs_fp64"] = round(forms[n]["value"] / forms["fp64"]["value"], 3)
    out["forms"] = forms
    print(json.dumps(out))


if __name__ == "__main__":
    main()
