"""Benchmark: 26-qubit complex128 state-vector bandwidth circuit on one MI355X.

Workload: n=26 state vector (2^26 amplitudes = 1 GiB), depth-20 circuit; each
layer is a Haar-random 2x2 unitary on every qubit, then CZ on (1,2),(3,4),...
in even layers and (2,3),(4,5),... in odd layers: 520 dense one-qubit gates and
250 CZ = 770 gates.  One step = one full pass of the circuit over the resident
state.  The state lives in HBM before the timed region starts.

Reported (one JSON line):
  value      gates/s of the fused plan (LDS-tile sweeps), the headline
  extra      gates/s and HBM GB/s with and without fusion, sweep counts
  roofline   the dominant kernel (k_tile, one fused sweep): algorithmic bytes
             per launch (2 x 16 x 2^26 B) / its average duration, from HIP
             events on the context stream around launches of that kernel alone
  cpu_baseline  the CPU oracle (cpu_ref, OpenMP) on a bounded sample of the
             same gates at the same size, on the host cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "HBM GB/s & gates/s at 26q c128 SV; batched adjoint grads/s with p99 latency"
HBM_PEAK_GBPS = 8000.0
N, DEPTH = 26, 20


FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = ROOT / "profiles" / "k_tile_26q_depth20_traffic.txt"


def read_traffic():
    """HBM bytes per k_tile launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (calibrated on k_gate1), from the committed profile summary."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def circuit_bytes(n, depth):
    """Algorithmic HBM bytes of the unfused circuit: a dense 1q gate reads and
    writes every amplitude (2*16*2^n B); CZ touches its |11> quarter (2*16*2^(n-2) B)."""
    n1q = n * depth
    ncz = sum(len(range(1 if l % 2 == 0 else 2, n, 2)) for l in range(depth))
    return n1q, ncz, n1q * 32 * 2 ** n + ncz * 32 * 2 ** (n - 2)


def time_program(ctx, prog, sv, steps, warmup, s0, s1):
    for _ in range(warmup):
        prog.run(sv.storage)
    ctx.sync()
    ctx.record(s0)
    for _ in range(steps):
        prog.run(sv.storage)
    ctx.record(s1)
    ctx.sync()
    return ctx.elapsed_ms(s0, s1)


def cpu_baseline(circ, n, budget_s=12.0, is_dm=False):
    """cpu_ref on the host cores: the circuit's instructions in order on a
    host vector (2^n amplitudes, or 4^n entries for a density matrix) until
    the time budget is spent; returns (instructions run, seconds, threads)."""
    import oracle
    from mossquantum_amd import lower
    lw = lower(circ, n, is_dm)
    psi = np.zeros(1 << (2 * n if is_dm else n), dtype=np.complex128)
    psi[0] = 1
    recs, pool = oracle._records(lw)
    lib = oracle.lib()
    t0 = time.perf_counter()
    done = 0
    one = (oracle.RefOp * 1)()
    while done < len(lw.records) and time.perf_counter() - t0 < budget_s:
        one[0] = recs[done]
        lib.ref_run_program(oracle._p(psi), n, int(is_dm), 1, one, oracle._p(pool))
        done += 1
    dt = time.perf_counter() - t0
    return done, dt, oracle.num_threads()


def cpu_baselines(circ_c1):
    """The CPU oracle (cpu_ref) on every configuration, on this host's cores
    and on one thread where stated; bounded samples, each named."""
    import oracle
    from mossquantum_amd import lower
    from mossquantum_amd.workloads import hea_angles, hea_circuit, tfim
    cores = oracle.num_threads()
    out = {"cores": cores, "kind": "port"}
    # C1 on one thread (the all-core figure is the headline's cpu_baseline)
    oracle.set_threads(1)
    try:
        done, dt, _ = cpu_baseline(circ_c1, N, budget_s=6.0)
        out["c1_26q_one_thread"] = {"gates_per_s": done / dt, "sample": f"first {done} gates, {dt:.1f} s"}
        # C0: value and full adjoint gradient, 12 qubits, 10 layers
        n, L = 12, 10
        th = hea_angles(n, L, 20260012)
        lw = lower(hea_circuit(n, L, th), n, False)
        t0 = time.perf_counter()
        oracle.adjoint(lw, tfim(n).groups())
        one = time.perf_counter() - t0
    finally:
        oracle.set_threads(cores)
    t0 = time.perf_counter()
    oracle.adjoint(lw, tfim(n).groups())
    out["c0_12q_adjoint_ms"] = {"one_thread": one * 1e3, "all_cores": (time.perf_counter() - t0) * 1e3}
    # C2: the 13-qubit noisy density matrix, instructions in order for a bounded time
    n, L = 13, 8
    c2 = hea_circuit(n, L, hea_angles(n, L, 20260213), noise=(0.01, 0.02))
    done, dt, _ = cpu_baseline(c2, n, budget_s=8.0, is_dm=True)
    nops = len(list(c2.flat()))
    out["c2_13q_density"] = {"instructions_per_s": done / dt, "full_run_s_at_that_rate": nops / (done / dt),
                             "sample": f"first {done} of {nops} instructions, {dt:.1f} s"}
    log(f"cpu baselines: C1 one thread {out['c1_26q_one_thread']['gates_per_s']:.1f} gates/s; C0 adjoint "
        f"{out['c0_12q_adjoint_ms']['one_thread']:.1f} / {out['c0_12q_adjoint_ms']['all_cores']:.1f} ms; "
        f"C2 {out['c2_13q_density']['instructions_per_s']:.1f} instructions/s")
    return out


def batched_bytes(n, layers=6):
    """Algorithmic bytes of one pure-state adjoint (one sweep per gate): forward
    (3n-1) gates per layer, reverse two vectors per gate, plus the observe pass."""
    gates = layers * (3 * n - 1)
    return 32 * 2 ** n * (3 * gates + 1.5 * (n - 1))


def bench_batched(ctx, B=512, layers=6, seed=20260300, cpu=True):
    """512 independent states, n ~ U{10..20}, 6-layer RY/RZ + CNOT-ladder ansatz
    with own seeded angles, XXZ loss; all gradients in one engine run.  Items
    of 12+ qubits run as batched tile sweeps; the gate-by-gate engine over the
    same batch is timed beside it."""
    import ctypes as C
    import mossquantum_amd as mq
    from mossquantum_amd import _lib
    from mossquantum_amd.workloads import c3_batch, hea_circuit, xxz
    sizes, circs, obs = c3_batch(B, layers, seed)
    rng = np.random.default_rng(seed + 2)
    dobs = [mq.DeviceObservable(o, n, ctx) for o, n in zip(obs, sizes)]
    oh = (C.c_void_p * B)(*[o.h.value for o in dobs])
    out = {"B": B, "layers": layers, "n_range": [10, 20]}
    results = {}
    for fuse in (True, False):
        t0 = time.perf_counter()
        progs = [mq.Program(c, n, False, ctx, fuse=fuse) for c, n in zip(circs, sizes)]
        log(f"batched: {B} programs lowered{' and planned' if fuse else ''} in {time.perf_counter() - t0:.1f} s")
        offs = np.zeros(B, dtype=np.int64)
        tot = 0
        for i, p in enumerate(progs):
            offs[i] = tot
            tot += p.lowered.nderiv
        vals = np.zeros(B)
        grads = np.zeros(tot)
        ph = (C.c_void_p * B)(*[p.h.value for p in progs])

        def run_engine():
            _lib.check(ctx.lib.mossq_batch_adjoint(ctx.h, B, ph, oh, offs.ctypes.data_as(C.POINTER(C.c_int64)),
                                                   _lib.dptr(vals), _lib.dptr(grads)))

        run_engine()  # warm-up (allocations, code objects)
        reps = 3 if fuse else 1
        ctx.sync()
        w0 = time.perf_counter()
        ctx.record(6)
        for _ in range(reps):
            run_engine()
        ctx.record(7)
        wall = (time.perf_counter() - w0) / reps
        dev_ms = ctx.elapsed_ms(6, 7) / reps
        key = "tiled" if fuse else "stepped"
        log(f"batched engine ({key}): {wall * 1e3:.1f} ms wall, {dev_ms:.1f} ms device per batch")
        results[key] = (wall, dev_ms, vals.copy(), grads.copy())
        if fuse:
            # plain per-circuit path: one mossq_adjoint per item (fused adjoint for 12+ qubits)
            pv = np.zeros(B)
            pg = np.zeros(tot)
            w0 = time.perf_counter()
            for i, (p, o) in enumerate(zip(progs, dobs)):
                v = C.c_double(0)
                g = np.zeros(max(p.lowered.nderiv, 1))
                _lib.check(ctx.lib.mossq_adjoint(ctx.h, p.h, o.h, None, 0, None, C.byref(v), _lib.dptr(g)))
                pv[i] = v.value
                pg[offs[i]:offs[i] + p.lowered.nderiv] = g[:p.lowered.nderiv]
            plain = time.perf_counter() - w0
            log(f"per-circuit path: {plain * 1e3:.1f} ms")
            out.update({"per_circuit_ms": plain * 1e3, "per_circuit_grads_per_s": B / plain,
                        "max_abs_grad_diff_vs_per_circuit": float(np.max(np.abs(grads - pg))),
                        "max_abs_value_diff_vs_per_circuit": float(np.max(np.abs(vals - pv)))})
            # last on these programs: it rebinds them to new angles
            out.update(bench_batch_graph(ctx, progs, ph, oh, offs, tot, vals, grads, sizes, layers, rng, reps))
        del progs
    wall, dev_ms, vals, grads = results["tiled"]
    swall, sdev, svals, sgrads = results["stepped"]
    out.update({"ms_per_batch": wall * 1e3, "device_ms": dev_ms, "grads_per_s": B / wall,
                "param_derivs_per_s": grads.size / wall, "speedup_vs_per_circuit": out["per_circuit_ms"] / 1e3 / wall,
                "stepped_engine_ms": swall * 1e3, "stepped_engine_grads_per_s": B / swall,
                "max_abs_grad_diff_tiled_vs_stepped": float(np.max(np.abs(grads - sgrads))),
                "max_abs_value_diff_tiled_vs_stepped": float(np.max(np.abs(vals - svals))),
                "unfused_convention_gbps": sum(batched_bytes(n, layers) for n in sizes) / wall / 1e9})
    if cpu:
        import oracle
        from mossquantum_amd import lower
        # one state of every size, timed with the oracle; the batch estimate sums per-size times
        per_size = {}
        for n in range(10, 21):
            c = hea_circuit(n, layers, rng.uniform(0, 2 * np.pi, 2 * n * layers))
            lw = lower(c, n, False)
            t1 = time.perf_counter()
            oracle.adjoint(lw, xxz(n).groups())
            per_size[n] = time.perf_counter() - t1
        est = sum(per_size[n] for n in sizes)
        out["cpu_estimate_s"] = est
        out["cpu_grads_per_s"] = B / est
        out["cpu_per_size_s"] = per_size
        out["cpu_sample"] = ("one adjoint per size n=10..20 timed with cpu_ref (OpenMP, all cores); the batch "
                             "time sums the per-size times over the batch's sizes")
        log(f"cpu estimate: {est:.1f} s for the batch")
    return out


REBIND_THREADS = 16  # host threads that rebind programs at once


def bench_batch_graph(ctx, progs, ph, oh, offs, tot, vals, grads, sizes, layers, rng, reps):
    """The batch captured once as a HIP graph (mossq_batch_capture) and
    replayed (mossq_batch_graph_run): replays at the same angles must equal
    the engine call bit for bit; then every program is rebound to new angles
    (mossq_program_update, lowering done beforehand) and the graph replayed,
    which must equal an engine call on the rebound programs."""
    import ctypes as C
    from mossquantum_amd import _lib, lower
    from mossquantum_amd.workloads import hea_circuit
    B = len(progs)
    optr = offs.ctypes.data_as(C.POINTER(C.c_int64))
    gh = C.c_void_p()
    t0 = time.perf_counter()
    _lib.check(ctx.lib.mossq_batch_capture(ctx.h, B, ph, oh, optr, C.byref(gh)))
    cap_ms = (time.perf_counter() - t0) * 1e3
    gv, gg = np.zeros(B), np.zeros(max(tot, 1))

    def replay():
        _lib.check(ctx.lib.mossq_batch_graph_run(ctx.h, gh, _lib.dptr(gv), _lib.dptr(gg)))

    try:
        replay()  # warm-up
        ctx.sync()
        w0 = time.perf_counter()
        ctx.record(6)
        for _ in range(reps):
            replay()
        ctx.record(7)
        wall = (time.perf_counter() - w0) / reps
        dev = ctx.elapsed_ms(6, 7) / reps
        same = bool(np.array_equal(gv, vals) and np.array_equal(gg[:tot], grads[:tot]))
        log(f"batch graph: capture {cap_ms:.1f} ms, replay {wall * 1e3:.1f} ms wall, {dev:.1f} ms device; "
            f"equal to the engine call: {same}")
        # rebinding to new angles: host lowering first (not timed), then updates + one replay
        new = []
        for p, n in zip(progs, sizes):
            lw = lower(hea_circuit(n, layers, rng.uniform(0, 2 * np.pi, 2 * n * layers)), n, False)
            recs, pool = lw.arrays()
            new.append((p, lw, recs, pool))
        # one library call: replanning on host threads, one upload of all values
        ph = (C.c_void_p * B)(*[p.h for p, _, _, _ in new])
        nops = (C.c_int * B)(*[len(lw.records) for _, lw, _, _ in new])
        rp = (C.c_void_p * B)(*[C.addressof(r) for _, _, r, _ in new])
        pp = (C.c_void_p * B)(*[pl.ctypes.data for _, _, _, pl in new])
        lens = (C.c_uint64 * B)(*[pl.size for _, _, _, pl in new])
        w0 = time.perf_counter()
        _lib.check(ctx.lib.mossq_program_update_many(ctx.h, B, ph, nops, rp, pp, lens, REBIND_THREADS, None))
        upd = time.perf_counter() - w0
        replay()
        rebind = time.perf_counter() - w0
        for p, lw, _, _ in new:
            p.lowered = lw
        ev, eg = np.zeros(B), np.zeros(max(tot, 1))
        _lib.check(ctx.lib.mossq_batch_adjoint(ctx.h, B, ph, oh, optr, _lib.dptr(ev), _lib.dptr(eg)))
        same_new = bool(np.array_equal(gv, ev) and np.array_equal(gg, eg))
        log(f"rebind {B} programs: {upd * 1e3:.1f} ms, + replay {rebind * 1e3:.1f} ms; "
            f"equal to the engine call: {same_new}")
        w0 = time.perf_counter()  # the same rebinding on one host thread
        _lib.check(ctx.lib.mossq_program_update_many(ctx.h, B, ph, nops, rp, pp, lens, 1, None))
        upd1 = time.perf_counter() - w0
        log(f"rebind on one host thread: {upd1 * 1e3:.1f} ms")
    finally:
        _lib.check(ctx.lib.mossq_graph_destroy(gh))
    return {"graph_capture_ms": cap_ms, "graph_ms_per_batch": wall * 1e3, "graph_device_ms": dev,
            "graph_grads_per_s": B / wall, "graph_equals_engine": same, "rebind_update_ms": upd * 1e3,
            "rebind_threads": REBIND_THREADS, "rebind_update_1thread_ms": upd1 * 1e3,
            "rebind_and_replay_ms": rebind * 1e3, "rebind_replay_equals_engine": same_new}


def bench_density(ctx, n=13, layers=8, seed=20260213, reps=2):
    """13-qubit density matrix (4^13 entries = 1 GiB), the ansatz with 8 layers,
    Depolarizing(0.01) then AmplitudeDamping(0.02) on every touched qubit after
    every gate; TFIM energy, trace and purity; fused and unfused plans."""
    import mossquantum_amd as mq
    from mossquantum_amd.workloads import hea_angles, hea_circuit, tfim
    c = hea_circuit(n, layers, hea_angles(n, layers, seed), noise=(0.01, 0.02))
    nops = len(list(c.flat()))
    out = {"n": n, "layers": layers, "instructions": nops, "state_bytes": 16 * 4 ** n}
    states = {}
    for fuse in (True, False):
        prog = mq.Program(c, n, True, ctx, fuse=fuse)
        dm = mq.DensityMatrix.zeros(n, ctx)
        prog.run(dm.storage)  # warm-up pass (not timed)
        ms = []
        for _ in range(reps):
            z = mq.DensityMatrix.zeros(n, ctx)
            ctx.sync()
            ctx.record(8)
            prog.run(z.storage)
            ctx.record(9)
            ms.append(ctx.elapsed_ms(8, 9))
            dm = z
        key = "fused" if fuse else "unfused"
        t = min(ms)
        sweeps = prog.sweeps()
        out[key] = {"ms": t, "sweeps": sweeps, "instructions_per_s": nops / (t / 1e3),
                    "hbm_gbps": sweeps * 2 * 16 * 4 ** n / (t / 1e3) / 1e9,
                    "effective_gbps_unfused_bytes": nops * 2 * 16 * 4 ** n / (t / 1e3) / 1e9}
        states[key] = dm
        log(f"density {key}: {t:.1f} ms, {sweeps} sweeps")
    dm = states["fused"]
    out["energy"] = mq.expectation(tfim(n), dm).real
    out["trace"] = mq.trace(dm).real
    out["purity"] = mq.purity(dm)
    out["energy_unfused"] = mq.expectation(tfim(n), states["unfused"]).real
    out["fused_vs_unfused_energy_diff"] = abs(out["energy"] - out["energy_unfused"])
    return out


def bench_c0(ctx, n=12, layers=10, seed=20260012, reps=50):
    """The reference's CPU-runnable case: 12 qubits, 10 ansatz layers (350
    gates, 240 parameters), TFIM energy.  A latency case: forward run eager and
    as a captured HIP graph replay, the adjoint gradient, and the
    parameter-shift gradient (481 evaluations in one batched engine run)."""
    import mossquantum_amd as mq
    from mossquantum_amd.workloads import hea_angles, hea_circuit, tfim
    th = hea_angles(n, layers, seed)
    tpl = hea_circuit(n, layers)
    c = tpl.bind(th)
    out = {"n": n, "layers": layers, "gates": len(list(c.flat())), "parameters": len(th)}
    for fuse in (True, False):
        prog = mq.Program(c, n, False, ctx, fuse=fuse)
        sv = mq.StateVector.zeros(n, ctx)
        graph = prog.capture(sv.storage)
        for run in (lambda: prog.run(sv.storage), graph.launch):
            run()
        ctx.sync()
        key = "fused" if fuse else "unfused"
        res = {"sweeps": prog.sweeps()}
        for name, run in (("eager", lambda: prog.run(sv.storage)), ("graph", graph.launch)):
            ctx.sync()
            t0 = time.perf_counter()
            ctx.record(8)
            for _ in range(reps):
                run()
            ctx.record(9)
            ctx.sync()
            res[f"{name}_us_per_run"] = (time.perf_counter() - t0) / reps * 1e6
            res[f"{name}_device_us_per_run"] = ctx.elapsed_ms(8, 9) * 1e3 / reps
        out[key] = res
        log(f"c0 forward {key}: eager {res['eager_us_per_run']:.1f} us, graph {res['graph_us_per_run']:.1f} us "
            f"per run ({res['sweeps']} sweeps)")
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.zeros(n, ctx))
    mq.value_and_gradient(mq.AdjointGradient(), obj, th + 0.5)  # warm-up at other angles
    # a call at new angles, as in an optimiser loop: the objective's program is
    # rebound in place (its plans kept) and differentiated
    t0 = time.perf_counter()
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    out["adjoint_ms"] = (time.perf_counter() - t0) * 1e3
    out["energy"] = v
    est = mq.estimate_gradient(mq.ParameterShift(), obj, th)
    t0 = time.perf_counter()
    est = mq.estimate_gradient(mq.ParameterShift(), obj, th)
    out["parameter_shift_ms"] = (time.perf_counter() - t0) * 1e3
    out["parameter_shift_evaluations"] = est.evaluations
    out["parameter_shift_vs_adjoint_max_abs_diff"] = float(np.max(np.abs(est.gradient - g)))
    log(f"c0 adjoint {out['adjoint_ms']:.2f} ms; parameter shift {out['parameter_shift_ms']:.1f} ms "
        f"({est.evaluations} evaluations)")
    return out


def bench_density_adjoint(ctx, n=12, layers=8, seed=20260412):
    """Value and full gradient of the TFIM energy of a 12-qubit noisy density
    matrix (the serving mix's largest density request: 8 layers, 192
    parameters, 1016 instructions), fused chunk plans against the step-by-step
    reverse pass; device time from HIP events around the library call."""
    import ctypes as C
    import mossquantum_amd as mq
    from mossquantum_amd import _lib
    from mossquantum_amd.backend import DeviceObservable, Program
    from mossquantum_amd.workloads import hea_angles, hea_circuit, tfim
    c = hea_circuit(n, layers, hea_angles(n, layers, seed), noise=(0.01, 0.02))
    lw = mq.lower(c, n, True)
    dob = DeviceObservable(tfim(n), n, ctx)
    out = {"n": n, "layers": layers, "instructions": len(lw.records), "parameters": lw.nderiv}
    grads = {}
    for fuse in (True, False):
        prog = Program(None, n, True, ctx, fuse=fuse, lowered=lw)
        g = np.zeros(lw.nderiv)
        v = C.c_double()
        st = mq.DensityMatrix.zeros(n, ctx)
        ms = []
        for _ in range(2):
            ctx.sync()
            ctx.record(8)
            _lib.check(ctx.lib.mossq_adjoint(ctx.h, prog.h, dob.h, st.storage.ptr, 1, None, C.byref(v), _lib.dptr(g)))
            ctx.record(9)
            ms.append(ctx.elapsed_ms(8, 9))
        key = "fused" if fuse else "step_by_step"
        out[key] = {"ms": min(ms), "value": v.value}
        grads[key] = g.copy()
        log(f"density adjoint {key}: {min(ms):.1f} ms")
    out["speedup"] = out["step_by_step"]["ms"] / out["fused"]["ms"]
    out["max_abs_grad_diff"] = float(np.max(np.abs(grads["fused"] - grads["step_by_step"])))
    return out


def serving_mix(ctx, nreq, rng, tpls):
    """`nreq` requests of the serving mix: 70% pure states n ~ U{8..24}, 30%
    density matrices n ~ U{6..12} with the noisy channels, layers ~ U{4..8},
    seeded angles; [(template, pool, angles, layers)]."""
    from mossquantum_amd.serving import RequestTemplate
    from mossquantum_amd.workloads import hea_circuit, tfim

    def template(kind, n, layers):
        key = (kind, n, layers)
        if key not in tpls:
            noise = (0.01, 0.02) if kind == "dm" else None
            tpls[key] = RequestTemplate(hea_circuit(n, layers, noise=noise), n, kind == "dm", tfim(n), ctx)
        return tpls[key]

    reqs = []
    for _ in range(nreq):
        dm = rng.random() < 0.3
        n = int(rng.integers(6, 13)) if dm else int(rng.integers(8, 25))
        layers = int(rng.integers(4, 9))
        tpl = template("dm" if dm else "sv", n, layers)
        th = rng.uniform(0, 2 * np.pi, tpl.nparams)
        reqs.append((tpl, lambda tpl=tpl, th=th: tpl.bind(th), th, layers))  # pools bound when submitted
    return reqs


def per_circuit_path(ctx, reqs):
    """The plain per-circuit GPU path on the same requests, one after the
    other: lower, plan (fused) and mossq_adjoint each request; req/s."""
    import ctypes as C
    from mossquantum_amd import _lib
    from mossquantum_amd.backend import Program
    t0 = time.perf_counter()
    for tpl, pool, _, _ in reqs:
        prog = Program(None, tpl.n, tpl.is_density, ctx, fuse=True, lowered=tpl.lowered, pool=pool())
        v = C.c_double()
        g = np.zeros(max(tpl.lowered.nderiv, 1))
        _lib.check(ctx.lib.mossq_adjoint(ctx.h, prog.h, tpl.obs.h, None, 0, None, C.byref(v), _lib.dptr(g)))
        del prog
    dt = time.perf_counter() - t0
    return {"requests": len(reqs), "seconds": dt, "throughput_req_s": len(reqs) / dt}


def cpu_serving(reqs, budget_s=40.0):
    """cpu_ref's serial per-request loop (value and adjoint gradient) on the
    mix: for every (kind, qubits) class in the sample one request is timed
    (smallest first, until the budget is spent), and each request counts that
    time scaled by its layer count (the work is linear in the layers).
    Density requests above 10 qubits are not run: the reference's checkpoint
    before every channel needs 4^n x 16 B x the channel count of host memory
    (about 64 GiB at 11 qubits, 256 GiB at 12), so their times are scaled
    from the 10-qubit class by 4^(n-10) and the instruction count."""
    import oracle
    from mossquantum_amd import lower
    from mossquantum_amd.workloads import hea_circuit, tfim
    classes = {}
    for tpl, _, th, layers in reqs:
        classes.setdefault((tpl.is_density, tpl.n), (th, layers))
    times, t_all = {}, time.perf_counter()
    for key in sorted(classes, key=lambda k: (k[1], k[0])):
        dm, n = key
        if (dm and n > 10) or time.perf_counter() - t_all > budget_s:
            continue
        th, layers = classes[key]
        lw = lower(hea_circuit(n, layers, th, noise=(0.01, 0.02) if dm else None), n, dm)
        t0 = time.perf_counter()
        oracle.adjoint(lw, tfim(n).groups())
        times[key] = (time.perf_counter() - t0) / layers  # per layer
    total, scaled, missing = 0.0, 0, 0
    for tpl, _, _, layers in reqs:
        key = (tpl.is_density, tpl.n)
        if key in times:
            total += times[key] * layers
        elif tpl.is_density and (True, 10) in times:
            inst = lambda m: 11 * m - 5  # noisy ansatz instructions per layer
            total += times[(True, 10)] * layers * 4 ** (tpl.n - 10) * inst(tpl.n) / inst(10)
            scaled += 1
        else:
            missing += 1
    counted = len(reqs) - missing
    return {"throughput_req_s": counted / total if total else None, "requests_counted": counted,
            "classes_timed": len(times), "density_scaled_from_10q": scaled, "cores": oracle.num_threads(),
            "sample": f"{len(reqs)} requests of the mix; one oracle adjoint per (kind, qubits) class timed "
                      f"and scaled by layers ({time.perf_counter() - t_all:.1f} s)"}


def bench_serving(ctx, rates=(50.0, 200.0, 2000.0), nreq=300, seed=20260400, budget=200 << 30, cpu=True):
    """Poisson request stream: 70% pure states n ~ U{8..24}, 30% density
    matrices n ~ U{6..12} with the noisy channels; 4-8 layers; TFIM energy +
    adjoint gradient; 200 GB admission budget.  `nreq` requests per rate; the
    plain per-circuit GPU path and the CPU oracle's serial loop run on a
    sample of the same mix."""
    from mossquantum_amd.serving import Server, serve_poisson
    rng = np.random.default_rng(seed)
    tpls = {}
    srv = Server(ctx, budget)
    out = {"requests_per_rate": nreq, "mix": "70% SV n~U{8..24}, 30% DM n~U{6..12}, layers~U{4..8}",
           "budget_bytes": budget, "rates": []}
    for rate in rates:
        reqs = serving_mix(ctx, nreq, rng, tpls)
        log(f"serving at {rate:.0f} req/s: {nreq} requests")
        _, stats = serve_poisson(srv, [(t, p) for t, p, _, _ in reqs], rate, seed=int(rate), progress=log)
        log(f"serving at {rate:.0f} req/s: {stats['throughput_req_s']:.1f} req/s, "
            f"p50 {stats['p50_ms']:.1f} ms, p99 {stats['p99_ms']:.1f} ms")
        out["rates"].append(stats)
    out["peak_device_bytes"] = srv.memory()[1]
    srv.close()
    sample = serving_mix(ctx, min(nreq, 300), np.random.default_rng(seed + 1), tpls)
    out["per_circuit_gpu"] = per_circuit_path(ctx, sample)
    log(f"serving mix through the per-circuit GPU path: {out['per_circuit_gpu']['throughput_req_s']:.1f} req/s")
    if cpu:
        out["cpu_serial"] = cpu_serving(sample)
        log(f"serving mix through the CPU oracle, serially: {out['cpu_serial']['throughput_req_s']:.2f} req/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-unfused", action="store_true", help="skip the unfused leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-adjoint leg")
    ap.add_argument("--no-density", action="store_true", help="skip the 13-qubit density-matrix leg")
    ap.add_argument("--serving-requests", type=int, default=300,
                    help="requests per arrival rate in the serving leg (0 skips it)")
    ap.add_argument("--serving-rates", default="50,200,2000",
                    help="arrival rates (req/s) of the serving leg, comma separated")
    ap.add_argument("--no-autotune", action="store_true", help="keep the planner's cost-model plan for the headline")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the fused 26-qubit headline (every other leg skipped), e.g. under a profiler")
    ap.add_argument("--leg", choices=("c0", "batched", "density", "density-adjoint", "serving"),
                    help="run only this secondary leg and print its JSON (for profiling one workload)")
    args = ap.parse_args()
    if args.headline_only:
        args.no_cpu = args.no_unfused = args.no_batched = args.no_density = True
        args.serving_requests = 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (items and requests would shard "
              "across GPUs without exchange, a single large state by its top qubits); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import mossquantum_amd as mq
    from mossquantum_amd.workloads import haar_layers, random_circuit

    rates = tuple(float(r) for r in args.serving_rates.split(","))
    if args.leg:
        ctx = mq.default_context()
        legs = {"c0": lambda: bench_c0(ctx), "batched": lambda: bench_batched(ctx, cpu=not args.no_cpu),
                "density": lambda: bench_density(ctx), "density-adjoint": lambda: bench_density_adjoint(ctx),
                "serving": lambda: bench_serving(ctx, rates=rates, nreq=args.serving_requests,
                                                 cpu=not args.no_cpu)}
        print(json.dumps({"leg": args.leg, "result": legs[args.leg]()}))
        return

    gold = ROOT / "tests" / "golden" / "c1_haar26.npz"
    U = np.load(gold)["unitaries"] if gold.exists() else haar_layers(N, DEPTH, 20260126)
    circ = random_circuit(N, DEPTH, U)
    n1q, ncz, alg_bytes = circuit_bytes(N, DEPTH)
    ngates = n1q + ncz
    ctx = mq.default_context()
    log(f"workload n={N} depth={DEPTH} gates={ngates} algorithmic bytes/step={alg_bytes/1e9:.1f} GB")

    sv = mq.StateVector.zeros(N, ctx)
    fused = mq.Program(circ, N, False, ctx, fuse=True)
    autotune = None
    if not args.no_autotune:
        # plan selection by measurement (outside the timed region): every
        # planner variant timed on a scratch state, the fastest kept
        model_ms = time_program(ctx, fused, sv, 2, 1, 0, 1) / 2
        variants, tuned_ms = fused.autotune(reps=2)
        autotune = {"variants_timed": variants, "model_choice_ms": model_ms, "tuned_ms": tuned_ms}
        log(f"autotune: {variants} plan variants, model choice {model_ms:.3f} ms, kept {tuned_ms:.3f} ms")
        sv = mq.StateVector.zeros(N, ctx)
    nsweeps = fused.sweeps()
    log(f"fused plan: {nsweeps} sweeps per step")
    ms_f = time_program(ctx, fused, sv, args.steps, args.warmup, 0, 1)
    step_f = ms_f / args.steps
    sweep_bytes = nsweeps * 32 * 2 ** N
    log(f"fused: {step_f:.3f} ms/step")

    extra = {"fused": {"ms_per_step": step_f, "gates_per_s": ngates / (step_f / 1e3), "sweeps": nsweeps,
                       "hbm_gbps": sweep_bytes / (step_f / 1e3) / 1e9,
                       "effective_gbps_unfused_bytes": alg_bytes / (step_f / 1e3) / 1e9,
                       "autotune": autotune}}
    # roofline of the dominant kernel: every fused launch is one k_tile sweep
    tile_avg_ms = step_f / nsweeps
    tile_bytes = 32 * 2 ** N
    achieved = tile_bytes / (tile_avg_ms / 1e3) / 1e9
    # algorithmic flops: a dense 2x2 complex update is 28 flops per amplitude pair
    tile_flops = n1q * 2 ** (N - 1) * 28 / nsweeps
    tile_tflops = tile_flops / (tile_avg_ms / 1e3) / 1e12

    if not args.no_unfused:
        sv_u = mq.StateVector.zeros(N, ctx)
        unf = mq.Program(circ, N, False, ctx, fuse=False)
        usteps = max(1, min(args.steps, 3))
        ms_u = time_program(ctx, unf, sv_u, usteps, 1, 2, 3)
        step_u = ms_u / usteps
        log(f"unfused: {step_u:.3f} ms/step")
        # parity of the two plans after the same number of passes
        for _ in range(args.steps + args.warmup - (usteps + 1)):
            unf.run(sv_u.storage)
        ctx.sync()
        a, b = sv.to_host(), sv_u.to_host()
        extra["fused_vs_unfused_max_abs_diff"] = float(np.max(np.abs(a - b)))
        del a, b
        # one-qubit dense sweeps alone (k_gate1): the first layer's 26 gates, on their own state
        from mossquantum_amd import QCircuit
        layer = QCircuit(list(circ.flat())[:N])
        g1 = mq.Program(layer, N, False, ctx, fuse=False)
        sv_g = mq.StateVector.zeros(N, ctx)
        ms_g1 = time_program(ctx, g1, sv_g, 4, 1, 4, 5)
        del sv_g
        g1_avg = ms_g1 / (4 * N)
        extra["unfused"] = {"ms_per_step": step_u, "gates_per_s": ngates / (step_u / 1e3),
                            "hbm_gbps": alg_bytes / (step_u / 1e3) / 1e9,
                            "gate1_avg_ms": g1_avg, "gate1_gbps": tile_bytes / (g1_avg / 1e3) / 1e9}
        del sv_u
        # the reference's own fusion (fuse_gates: adjacent gates merged while
        # their support stays within 5 qubits), then one sweep per merged gate
        fg = mq.fuse_gates(circ, 5)
        kprog = mq.Program(fg, N, False, ctx, fuse=False)
        sv_k = mq.StateVector.zeros(N, ctx)
        ms_k = time_program(ctx, kprog, sv_k, 1, 0, 6, 7)
        sv_1 = mq.StateVector.zeros(N, ctx)
        fused.run(sv_1.storage)
        ctx.sync()
        a, b = sv_1.to_host(), sv_k.to_host()
        extra["fuse_gates_k5"] = {"ms_per_step": ms_k, "gates_per_s": ngates / (ms_k / 1e3),
                                  "merged_gates": len(list(fg.flat())),
                                  "max_abs_diff_vs_fused": float(np.max(np.abs(a - b)))}
        log(f"fuse_gates(5): {len(list(fg.flat()))} gates, {ms_k:.1f} ms/step")
        del a, b, sv_k, sv_1
    norm = mq.dot(sv, sv).real
    extra["final_norm_minus_one"] = norm - 1.0
    del sv
    if not args.headline_only:
        extra["c0_hea12"] = bench_c0(ctx)
    if not args.no_batched:
        extra["batched_adjoint"] = bench_batched(ctx, cpu=not args.no_cpu)
    if not args.no_density:
        extra["density_13q"] = bench_density(ctx)
        extra["density_adjoint_12q"] = bench_density_adjoint(ctx)
    if args.serving_requests > 0:
        extra["serving"] = bench_serving(ctx, rates=rates, nreq=args.serving_requests, cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        done, dt, cores = cpu_baseline(circ, N)
        cpu = {"value": done / dt, "unit": "gates/s", "cores": cores, "kind": "port",
               "sample": f"first {done} of the {ngates} gates of this circuit at n=26 "
                         f"({dt:.1f} s, cpu_ref OpenMP, one sweep per gate)"}
        log(f"cpu baseline: {done} gates in {dt:.1f} s on {cores} threads")
        extra["cpu_baselines"] = cpu_baselines(circ)

    out = {
        "metric": METRIC,
        "value": ngates / (step_f / 1e3),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_f,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded Haar unitaries, state |0...0>)",
        "config": {"workload": "26q c128 SV random circuit, depth 20 (Haar 1q + alternating CZ), fused plan",
                   "n_qubits": N, "depth": DEPTH, "gates": ngates, "state_bytes": 16 * 2 ** N},
        "roofline": {"bound": "fp64-valu", "achieved": tile_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tile_tflops / FP64_PEAK_TFLOPS, "traffic": read_traffic(), "kernel": "k_tile",
                     "note": "FP64 FMA on the vector ALUs, no MFMA (MI355X's fp64 peak, 78.6 TFLOP/s, is the "
                             "same for the vector and the matrix units); a fused sweep carries ~40 dense gates "
                             "per amplitude pass (~17 flop/B), past the fp64 ridge (~10 flop/B); the HBM view "
                             "is hbm_achieved_gbps / hbm_frac",
                     "avg_launch_ms": tile_avg_ms, "flops_per_launch": tile_flops,
                     "bytes_per_launch": tile_bytes, "hbm_achieved_gbps": achieved,
                     "hbm_frac": achieved / HBM_PEAK_GBPS},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
