"""Benchmark of the GCMC energy engine on one MI355X.

Headline (the timed step): one batched Widom call of the stated Widom workload
-- 100,000 CBMC test insertions of rigid CO2 (10 trial positions x 10
orientations each) into a synthetic 1,536-atom framework in a 24 A cubic box,
empty box, Ewald alpha 0.265 / kmax 7, in the engine's fast form (FP64 pair
arithmetic with an FP32-seeded reciprocal and a tabulated erfc(ar)/r, an FP32
overlap screen, FP64 Ewald).  value = insertions/s over K timed steps with all
inputs resident in HBM; a step is one replay of the batch captured as a HIP
graph (graspa_widom_graph_launch), the same batches as queued enqueues, which
are timed beside it (--no-graph makes them the headline).

Also reported on the same line: the plain FP64 form of the same batch, a
lower-density variant of the framework (750 atoms) where most constructions
complete, the plain FP64 single-box GCMC (translate:rotate:insert:delete
1:1:1:1) moves/s with p50/p99 per-cycle latency, the roofline of the dominant
kernel (the pair-energy kernel: algorithmic FLOP per launch / launch time
measured with HIP events on the box's stream), and the oracle (FP64 C, one
core) timed on a bounded sample as the CPU baseline.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector / matrix peak
FP64_PEAK_TFLOPS = 78.6


SCREEN_FLOP = 21.0       # difference 3 + image 12 + r^2 5 + test 1
PAIR_FLOP = 52.0         # + LJ 10 + Coulomb 19 (erfc as 12) + accumulate 2
PAIR_FLOP_UNCHARGED = 33.0


def widom_flops(screened, evaluated, M, n_active_k, n_ok, charged=True):
    """Algorithmic FLOP of a Widom batch as executed, per unit costs: 21 per
    (trial site, atom) pair tested by the overlap screen (distance and test),
    52 per pair with full energies (33 uncharged), 16 per (k-vector, site) and
    12 per k-vector of the Fourier term of each completed insertion.  The pair
    counts are the kernel's own work counters."""
    f = screened * SCREEN_FLOP + evaluated * (PAIR_FLOP if charged else PAIR_FLOP_UNCHARGED)
    if charged:
        f += n_ok * n_active_k * (16.0 * M + 12.0)
    return f


def reference_flops(n_sys, K, O, chain, M, n_active_k, n_total, n_fb, n_ok, charged=True):
    """FLOP the reference's algorithm spends on the same batch: full energies of
    all K first-bead trials, of the O orientations when the first bead is
    placed, and the Fourier term of completed insertions."""
    pair = PAIR_FLOP if charged else PAIR_FLOP_UNCHARGED
    f = n_total * K * n_sys * pair + n_fb * O * chain * n_sys * pair
    if charged:
        f += n_ok * n_active_k * (16.0 * M + 12.0)
    return f


def load_traffic(kernel_prefix):
    """HBM bytes per launch of the Widom trial-evaluation kernel from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy)."""
    p = os.path.join(ROOT, "profiles", "widom_eval_kernel_config1_100k.json")
    try:
        d = json.load(open(p))
    except (OSError, ValueError):
        return None, None
    if not d.get("kernel", "").startswith(kernel_prefix):
        return None, None
    return d.get("traffic_bytes"), d


def run_widom(system, precision, n_insert, steps, warmup, seed=2024, timing=True, autotune=False, queued=False,
              graph=False):
    """queued=False: one synchronous graspa_widom_batch per step (summary back
    on the host after each batch).  queued=True: the steps are queued with
    graspa_widom_enqueue and their summaries summed on the device, read once
    after the last step (graspa_widom_totals).  graph=True (with queued): the
    batch is captured once as a HIP graph before the timed region and each
    step is one replay (graspa_widom_graph_launch), the same batches."""
    from graspa_amd import engine
    # torch's first import takes about a second; done here so that the GPU
    # does not sit idle between the warmup and the timed batches (after such a
    # gap the pair kernel ran 13% slower for the next several milliseconds)
    _sync_torch()
    setup = {}
    t0 = time.perf_counter()
    box = engine.Box(system, precision=precision)
    box.sync()
    setup["box_create_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    box.init_ewald()
    box.sync()
    setup["structure_factors_ms"] = (time.perf_counter() - t0) * 1e3
    intra, selfe = engine.exclusion_energy(system, 1)
    excl = intra + selfe
    # the first batch builds the certain-overlap voxel maps of the inserted
    # site types on the host (kept for later batches)
    t0 = time.perf_counter()
    box.widom_batch(1, 64, seed=seed + 2, first_index=0, excl=excl, results=False)
    setup["first_batch_with_voxel_maps_ms"] = (time.perf_counter() - t0) * 1e3
    tune = None
    if autotune:
        # launch shape of the pair-energy kernel, picked by timing every shape
        # that fits on a batch of the same size (other draws), before the timed region
        default_shape = box.eval_shape()
        t0 = time.perf_counter()
        chosen, tried = box.autotune_widom(1, n_insert, seed=seed + 1, reps=2, excl=excl)
        setup["autotune_ms"] = (time.perf_counter() - t0) * 1e3
        # the built-in shape: per_cu 0 means as many workgroups per CU as are
        # resident, i.e. the largest per_cu tried for its waves and staging
        same = [(shp["per_cu"], ms) for shp, ms in tried
                if shp["waves"] == default_shape["waves"] and shp["framework_lds"] == default_shape["framework_lds"]]
        default_ms = max(same)[1] if same and default_shape["per_cu"] == 0 else \
            next((ms for pc, ms in same if pc == default_shape["per_cu"]), None)
        tune = dict(shape=chosen, shapes_tried=len(tried), best_ms=min(ms for _, ms in tried),
                    default_shape=default_shape, default_shape_ms=default_ms,
                    slowest_ms=max(ms for _, ms in tried))
    for w in range(warmup):
        box.widom_batch(1, n_insert, seed=seed, first_index=0, excl=excl, results=False)
    if graph:
        t0 = time.perf_counter()
        box.widom_graph(1, n_insert, seed=seed, first_index=n_insert, excl=excl)
        setup["graph_capture_ms"] = (time.perf_counter() - t0) * 1e3
    if timing:
        box.timing_enable("widom", True)
    box.sync()
    _sync_torch()
    box.widom_totals(reset=True)
    t0 = time.perf_counter()
    summ = None
    tot = dict(sum_w=0.0, n=0.0, n_first_bead=0.0, n_ok=0.0, sum_we=0.0, screened_pairs=0.0, evaluated_pairs=0.0)
    for k in range(steps):
        if queued and graph:
            box.widom_graph_launch(1)
            continue
        if queued:
            box.widom_enqueue(1, n_insert, seed=seed, first_index=(k + 1) * n_insert, excl=excl)
            continue
        _, summ = box.widom_batch(1, n_insert, seed=seed, first_index=(k + 1) * n_insert, excl=excl, results=False)
        for key in tot:
            tot[key] += getattr(summ, key)
    dev_tot = box.widom_totals(reset=True)
    box.sync()
    _sync_torch()
    t1 = time.perf_counter()
    if queued:
        for key in tot:
            tot[key] = getattr(dev_tot, key)
    else:
        # the device-side totals must agree with the per-batch summaries
        assert dev_tot.n == tot["n"] and dev_tot.evaluated_pairs == tot["evaluated_pairs"]
    kern_ms, launches = box.timing_read() if timing else (0.0, 0)
    if timing:
        box.timing_enable("widom", False)
    nvec, nact = box.nvec()
    box.close()
    return dict(seconds=t1 - t0, kern_ms=kern_ms, launches=launches, nact=nact, autotune=tune, setup=setup, **tot)


def _sync_torch():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def mixed_error(system, n):
    """max |dE| of the mixed form against the plain FP64 form on n insertions
    whose constructions agree (same first bead and orientation)."""
    import numpy as np
    from graspa_amd import engine
    out = {}
    intra, selfe = engine.exclusion_energy(system, 1)
    for prec in (engine.GRASPA_FP64, engine.GRASPA_MIXED):
        box = engine.Box(system, precision=prec)
        box.init_ewald()
        res, _ = box.widom_batch(1, n, seed=77, excl=intra + selfe)
        out[prec] = engine.widom_results_to_numpy(res)
        box.close()
    a, b = out[engine.GRASPA_FP64], out[engine.GRASPA_MIXED]
    same = (a["ok"] == 1) & (b["ok"] == 1) & (a["first_bead"] == b["first_bead"]) & (a["orientation"] == b["orientation"])
    if not same.any():
        return dict(n_compared=0, max_abs_dE=None, max_abs_dE_kJmol=None)
    err = np.abs(a["energy"][same] - b["energy"][same])
    return dict(n_compared=int(same.sum()), max_abs_dE=float(err.max()), max_abs_dE_kJmol=float(err.max() * 0.01),
                median_abs_dE_kJmol=float(np.median(err) * 0.01),
                selection_agreement=float(same.sum() / max(1, ((a["ok"] == 1) | (b["ok"] == 1)).sum())))


def run_gcmc(system, cycles):
    import numpy as np
    from graspa_amd import engine
    box = engine.Box(system, precision=engine.GRASPA_FP64, capacity=512)
    box.init_ewald()
    mc = engine.MonteCarlo(box)
    box.sync()
    t0 = time.perf_counter()
    n, _ = mc.cycles(0, cycles, production=True)
    t1 = time.perf_counter()
    ct = mc.cycle_times(cycles)
    e, nm, cnt, ww, wn = mc.stats()
    mc.close()
    box.close()
    return dict(moves=int(n), seconds=t1 - t0, moves_per_s=n / (t1 - t0), cycle_p50_ms=float(np.percentile(ct, 50) * 1e3),
                cycle_p99_ms=float(np.percentile(ct, 99) * 1e3), n_molecules=int(nm[0]),
                accepted=int(cnt[0][1] + cnt[0][3] + cnt[0][5] + cnt[0][7]))


def run_batch_gcmc(n_boxes, cycles, precision, warm_cycles=1, capacity=2048):
    """Config-3 style batch: n_boxes independent CO2/N2 boxes (frameworks of
    500-6000 atoms per unit cell, 18-40 A, T 250-350 K, P 1e3-1e7 Pa), all on
    the device in one persistent launch; moves/s over `cycles` cycles per box
    and the per-cycle latency distribution over every box's cycles."""
    import numpy as np
    from graspa_amd import engine, synth
    systems = [synth.config3_box(b) for b in range(n_boxes)]
    boxes = []
    for s in systems:
        bx = engine.Box(s, precision=precision, capacity=capacity)
        bx.init_ewald()
        boxes.append(bx)
    bt = engine.Batch(boxes, [3000 + b for b in range(n_boxes)], list(range(n_boxes)), cycle_cap=max(cycles, 1))
    if warm_cycles:
        bt.run(warm_cycles, production=False)
    k0, l0 = bt.kernel_time()
    t0 = time.perf_counter()
    moves = bt.run(cycles, production=True)
    t1 = time.perf_counter()
    k1, l1 = bt.kernel_time()
    lat = np.concatenate([bt.cycle_times(i, cycles) for i in range(n_boxes)])
    n_mol = np.array([bt.stats(i).n_mol for i in range(n_boxes)])
    atoms = np.array([len(s.fw_type) for s in systems])
    out = dict(boxes=n_boxes, cycles_per_box=cycles, moves=int(moves), seconds=t1 - t0,
               moves_per_s=moves / (t1 - t0), kernel_ms=k1 - k0, launches=int(l1 - l0),
               cycle_p50_ms=float(np.percentile(lat, 50) * 1e3), cycle_p99_ms=float(np.percentile(lat, 99) * 1e3),
               framework_atoms_min=int(atoms.min()), framework_atoms_max=int(atoms.max()),
               mean_molecules_per_box=float(n_mol.sum(1).mean()), max_molecules_per_box=int(n_mol.sum(1).max()))
    bt.close()
    for bx in boxes:
        bx.close()
    return out


def gcmc_de_error(systems, cycles, seeds, capacity=2048, trace_cap=4000):
    """|dE| of the fast form against the plain FP64 form on the same boxes and
    draws: both forms run `cycles` cycles through the device-resident driver
    with move traces; along each box's common sequence (the trajectories stay
    the same while every accept/reject decision agrees) the energy change of
    every accepted move is compared term by term."""
    from graspa_amd import engine
    keys = ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald", "tail")
    traces = {}
    for prec in (engine.GRASPA_FP64, engine.GRASPA_MIXED):
        boxes = []
        for s in systems:
            bx = engine.Box(s, precision=prec, capacity=capacity)
            bx.init_ewald()
            boxes.append(bx)
        bt = engine.Batch(boxes, list(seeds), list(range(len(systems))), trace_cap=trace_cap, cycle_cap=1)
        bt.run(cycles, production=True)
        traces[prec] = [bt.trace(i) for i in range(len(systems))]
        bt.close()
        for bx in boxes:
            bx.close()
    n_cmp, err, same, moves, n_p, err_p = 0, 0.0, 0, 0, 0, 0.0
    for s, ta, tb in zip(systems, traces[engine.GRASPA_FP64], traces[engine.GRASPA_MIXED]):
        whole = len(ta) == len(tb)
        for a, b in zip(ta, tb):
            if (a.move, a.comp, a.mol, a.accepted) != (b.move, b.comp, b.mol, b.accepted):
                whole = False
                break
            moves += 1
            if a.accepted:
                n_cmp += 1
                err = max(err, max(abs(getattr(a.de, k) - getattr(b.de, k)) for k in keys))
            if 1e-300 < a.pacc < 1e300 and 1e-300 < b.pacc < 1e300:
                # every evaluated move, accepted or not: |ln Pacc_fast - ln Pacc_FP64| / beta
                # is the energy error of the move (for CBMC moves a Boltzmann-weighted
                # mixture of its trials' errors)
                n_p += 1
                err_p = max(err_p, abs(math.log(b.pacc) - math.log(a.pacc)) / s.beta)
        same += int(whole)
    # no compared moves: no error figure (not 0.0)
    return dict(boxes=len(systems), cycles=cycles, moves_in_common=moves, accepted_moves_compared=n_cmp,
                max_abs_dE_kJmol=err * 0.01 if n_cmp else None, evaluated_moves_compared=n_p,
                max_abs_dlnPacc_over_beta_kJmol=err_p * 0.01 if n_p else None,
                identical_trajectories=same)


def run_batch_single(system, cycles, precision, warm_cycles=0, capacity=1024, seed=11):
    """One box through the device-resident driver (batch of one)."""
    import numpy as np
    from graspa_amd import engine
    bx = engine.Box(system, precision=precision, capacity=capacity)
    bx.init_ewald()
    bt = engine.Batch([bx], [seed], [0], cycle_cap=max(cycles, 1))
    if warm_cycles:
        bt.run(warm_cycles, production=False)
    t0 = time.perf_counter()
    moves = bt.run(cycles, production=True)
    t1 = time.perf_counter()
    lat = bt.cycle_times(0, cycles)
    st = bt.stats(0)
    out = dict(moves=int(moves), seconds=t1 - t0, moves_per_s=moves / (t1 - t0),
               cycle_p50_ms=float(np.percentile(lat, 50) * 1e3), cycle_p99_ms=float(np.percentile(lat, 99) * 1e3),
               n_molecules=[int(x) for x in st.n_mol])
    bt.close()
    bx.close()
    return out


# The reference's CO2-in-MFI run (Examples/CO2-MFI/output.txt:171-212): 5,000
# initialization cycles, 104,868 moves in 7.967520 s of work
REFERENCE_CO2_MFI = dict(moves=24719 + 25130 + 13735 + 13766 + 27518, seconds=7.967520, final_molecules=26)


def run_co2_mfi(n_seeds=16, cycles=5000):
    """The reference's CO2-MFI deck (MFI 2x2x2 with CIF charges, rigid CO2,
    298 K, 1e4 Pa, translation : rotation : reinsertion : swap 1:1:1:1) on the
    device-resident driver: one box, then n_seeds independent boxes in one
    launch, in both forms; moves/s, per-cycle latency, final loading and
    acceptance ratios beside the reference's own run."""
    import numpy as np
    from graspa_amd import deck, engine
    s = deck.load_system(os.path.join(ROOT, "tests", "golden", "co2_mfi.npz"))
    ref = REFERENCE_CO2_MFI
    out = dict(deck="Examples/CO2-MFI (tests/golden/co2_mfi.npz)", cycles=cycles,
               reference=dict(ref, moves_per_s=ref["moves"] / ref["seconds"]))

    def leg(n, prec):
        boxes = []
        for _ in range(n):
            b = engine.Box(s, precision=prec, capacity=256)
            b.init_ewald()
            boxes.append(b)
        bt = engine.Batch(boxes, [101 + i for i in range(n)], list(range(n)), cycle_cap=cycles)
        t0 = time.perf_counter()
        moves = bt.run(cycles, production=False)
        t1 = time.perf_counter()
        lat = np.concatenate([bt.cycle_times(i, cycles) for i in range(n)])
        st = [bt.stats(i) for i in range(n)]
        c = np.array([x.counters[0] for x in st], dtype=np.float64).sum(0)
        r = dict(boxes=n, moves=int(moves), seconds=t1 - t0, moves_per_s=moves / (t1 - t0),
                 cycle_p50_ms=float(np.percentile(lat, 50) * 1e3), cycle_p99_ms=float(np.percentile(lat, 99) * 1e3),
                 final_molecules_mean=float(np.mean([x.n_mol[0] for x in st])),
                 acceptance=dict(translation=c[1] / max(c[0], 1), rotation=c[3] / max(c[2], 1),
                                 insertion=c[5] / max(c[4], 1), deletion=c[7] / max(c[6], 1),
                                 reinsertion=c[10] / max(c[9], 1)))
        bt.close()
        for b in boxes:
            b.close()
        return r

    # the reference-order parity driver (host-driven, one move at a time, plain
    # FP64: the reference's own form), over the first 1,000 cycles
    out["parity_driver_fp64_1000_cycles"] = run_gcmc(s, 1000)
    out["parity_driver_fp64_1000_cycles"]["over_reference"] = (
        out["parity_driver_fp64_1000_cycles"]["moves_per_s"] / out["reference"]["moves_per_s"])
    for name, prec in (("fp64", engine.GRASPA_FP64), ("mixed", engine.GRASPA_MIXED)):
        one = leg(1, prec)
        many = leg(n_seeds, prec)
        one["over_reference"] = one["moves_per_s"] / out["reference"]["moves_per_s"]
        many["over_reference"] = many["moves_per_s"] / out["reference"]["moves_per_s"]
        out[name + "_one_box"] = one
        out[name + "_%d_boxes" % n_seeds] = many
    return out


def run_batch_steady(n_boxes, cycles, precision, window=100, capacity=2048):
    """The 256-box distribution run for its stated cycles from empty boxes:
    moves/s over the whole run and over its last half (steady loading), the
    per-cycle latency of every box's cycles in the last half, loadings and
    the spread of the boxes' busy time (the launch lasts as long as its
    slowest box)."""
    import numpy as np
    from graspa_amd import engine, synth
    systems = [synth.config3_box(b) for b in range(n_boxes)]
    boxes = []
    for s in systems:
        bx = engine.Box(s, precision=precision, capacity=capacity)
        bx.init_ewald()
        boxes.append(bx)
    bt = engine.Batch(boxes, [3000 + b for b in range(n_boxes)], list(range(n_boxes)), cycle_cap=window)
    done, moves, secs, rows = 0, 0, 0.0, []
    while done < cycles:
        n = min(window, cycles - done)
        k0, _ = bt.kernel_time()
        t0 = time.perf_counter()
        mv = bt.run(n, production=True)
        t1 = time.perf_counter()
        k1, _ = bt.kernel_time()
        done += n
        rows.append((done, mv, t1 - t0, k1 - k0))
    half = [r for r in rows if r[0] > cycles // 2]
    lat = np.concatenate([bt.cycle_times(i, window) for i in range(n_boxes)])
    nm = np.array([bt.stats(i).n_mol.sum() for i in range(n_boxes)])
    per_box = np.array([bt.stats(i).moves for i in range(n_boxes)])
    busy = bt.busy_seconds()
    # algorithmic work of the moves made (the reference's algorithm on the
    # same moves: every pair of the moved or trial sites with every framework
    # and guest atom, 52 FLOP each; 16 FLOP per (active k, site) and 12 per
    # active k of the Fourier change)
    flop = 0.0
    for i, sy in enumerate(systems):
        st = bt.stats(i)
        n_act = boxes[i].nvec()[1]
        nf = len(sy.fw_type)
        load = st.loading()
        ng = float(sum(load[c] * sy.components[c].molsize for c in range(len(sy.components))))
        for c, comp in enumerate(sy.components):
            M = comp.molsize
            cnt = st.counters[c]
            body = cnt[0] + cnt[2]                     # translations, rotations with a molecule
            swap = cnt[4] + cnt[6] + cnt[8]            # insertions, deletions, Widom
            flop += body * (2 * M * (nf + ng) * PAIR_FLOP + n_act * (16.0 * 2 * M + 12.0))
            flop += swap * ((sy.n_trial_pos + sy.n_trial_orient * (M - 1)) * (nf + ng) * PAIR_FLOP
                            + n_act * (16.0 * M + 12.0))
    kern_s = sum(r[3] for r in rows) * 1e-3
    achieved = flop / kern_s / 1e12
    out = dict(boxes=n_boxes, cycles_per_box=cycles, moves=int(sum(r[1] for r in rows)),
               seconds=float(sum(r[2] for r in rows)), kernel_ms=float(sum(r[3] for r in rows)),
               moves_per_s=sum(r[1] for r in rows) / sum(r[2] for r in rows),
               steady_moves_per_s=sum(r[1] for r in half) / sum(r[2] for r in half),
               steady_cycle_p50_ms=float(np.percentile(lat, 50) * 1e3),
               steady_cycle_p99_ms=float(np.percentile(lat, 99) * 1e3),
               mean_molecules_per_box=float(nm.mean()), max_molecules_per_box=int(nm.max()),
               moves_per_box_max=int(per_box.max()), moves_per_box_median=float(np.median(per_box)),
               busy_fraction=float(busy.sum() / (n_boxes * kern_s)),
               busiest_box_seconds=float(busy.max()), median_box_seconds=float(np.median(busy)),
               roofline=dict(bound="valu", unit="TFLOP/s", achieved=achieved, peak=FP64_PEAK_TFLOPS,
                             frac=achieved / FP64_PEAK_TFLOPS, algorithmic_flop=flop,
                             note="algorithmic FLOP of the moves made (every pair of the moved or trial "
                                  "sites with every atom) over the batch kernel's time; the launch lasts as "
                                  "long as its busiest box (busy_fraction = summed box time / boxes x kernel "
                                  "time)"))
    bt.close()
    for bx in boxes:
        bx.close()
    return out


def run_served(n_jobs, rate, slots, precision, cycles_lo=50, cycles_hi=5000, seed=4000, time_slice_us=0.0,
               moves_per_launch=256):
    """Served stream: Poisson arrivals of config-3 style jobs, continuous
    batching; per-cycle latency over every job's cycles and aggregate moves/s.
    time_slice_us > 0: every box runs about that long per launch at its own
    pace (graspa_batch_set_time_slice) instead of moves_per_launch moves."""
    import numpy as np
    from graspa_amd import serve
    jobs = serve.poisson_jobs(n_jobs, rate, seed=seed, cycles_lo=cycles_lo, cycles_hi=cycles_hi)
    srv = serve.Server(n_slots=slots, precision=precision, moves_per_launch=moves_per_launch, cycle_cap=cycles_hi,
                       time_slice_us=time_slice_us)
    done = srv.run(jobs)
    srv.close()
    lat = np.concatenate([j.cycle_seconds for j in done])
    moves = sum(j.stats.moves for j in done)
    wait = np.array([j.admitted - j.arrival for j in done])
    return dict(jobs=n_jobs, arrival_rate_per_s=rate, slots=slots, cycles_range=[cycles_lo, cycles_hi],
                time_slice_us=time_slice_us, moves_per_launch=moves_per_launch,
                peak_device_bytes=int(srv.peak_bytes),
                total_cycles=int(sum(j.cycles for j in done)), moves=int(moves), seconds=srv.elapsed,
                moves_per_s=moves / srv.elapsed, launches=srv.launches,
                cycle_p50_ms=float(np.percentile(lat, 50) * 1e3), cycle_p99_ms=float(np.percentile(lat, 99) * 1e3),
                admission_wait_p50_ms=float(np.percentile(wait, 50) * 1e3),
                admission_wait_p99_ms=float(np.percentile(wait, 99) * 1e3))


def host_cpu():
    """CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return dict(model=model, nproc=os.cpu_count(), usable_cores=usable,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def cpu_baseline(system, n_insert):
    """The FP64 oracle (C, -O2) on a bounded sample of the same workload: one
    thread (like-for-like with the reference, whose CPU energy code is
    single-threaded), and OpenMP over insertions on the cores this process
    may use (capped by OMP_NUM_THREADS when set)."""
    from oracle import binding as ob
    o = ob.OracleSystem(system)
    _, skf, ska, _, _, _ = o.ewald_total()
    host = host_cpu()
    cap = int(host["omp_num_threads"]) if (host["omp_num_threads"] or "").isdigit() else host["usable_cores"]
    n_all = max(1, min(host["usable_cores"] or 1, cap))
    ob.lib().or_set_threads(1)
    t0 = time.perf_counter()
    o.widom_batch(1, ska, skf, seed=2024, n_insert=n_insert)
    t1 = time.perf_counter()
    one = n_insert / (t1 - t0)
    used = ob.lib().or_set_threads(n_all)
    n_big = n_insert * used
    t2 = time.perf_counter()
    o.widom_batch(1, ska, skf, seed=2024, n_insert=n_big)
    t3 = time.perf_counter()
    ob.lib().or_set_threads(1)
    return dict(value=n_big / (t3 - t2), unit="insertions/s", cores=used, kind="port",
                sample="%d Widom insertions of the same workload, FP64 oracle in C (gcc -O2), OpenMP over "
                       "insertions on %d threads, %.1f s" % (n_big, used, t3 - t2),
                one_thread=dict(value=one, sample="%d insertions, one thread, %.1f s" % (n_insert, t1 - t0)),
                host=host)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--insertions", type=int, default=100000)
    ap.add_argument("--gcmc-cycles", type=int, default=2000)
    ap.add_argument("--cpu-sample", type=int, default=15000)
    ap.add_argument("--quick", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in launch shape of the pair kernel")
    ap.add_argument("--sync-steps", action="store_true",
                    help="time one synchronous graspa_widom_batch per step instead of queued batches")
    ap.add_argument("--no-graph", action="store_true",
                    help="time queued batches (graspa_widom_enqueue) instead of replays of a captured graph")
    ap.add_argument("--batch-boxes", type=int, default=256)
    ap.add_argument("--batch-cycles", type=int, default=20)
    ap.add_argument("--batch-warm-cycles", type=int, default=100)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run 3 plain device-to-device copies of BYTES (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: boxes and Widom batches shard across GPUs with no communication, "
              "and this build runs on a single MI355X (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        import torch
        a = torch.empty(args.copy_calibration, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    from graspa_amd import engine, synth

    s1 = synth.config1()
    K, O, M = s1.n_trial_pos, s1.n_trial_orient, s1.components[0].molsize
    n_fw = len(s1.fw_type)
    # the pair kernel's launch time (HIP events on the box's stream) over
    # queued batches; the headline steps are the same batches replayed from a
    # captured HIP graph (measured faster, same totals), unless --no-graph
    r = run_widom(s1, engine.GRASPA_MIXED, args.insertions, args.steps, args.warmup, autotune=not args.no_autotune,
                  queued=not args.sync_steps)
    n_total = args.insertions * args.steps
    use_graph = not (args.no_graph or args.sync_steps)
    rv = (run_widom(s1, engine.GRASPA_MIXED, args.insertions, args.steps, args.warmup,
                    autotune=not args.no_autotune, timing=False, queued=True, graph=True) if use_graph else r)
    value = n_total / rv["seconds"]
    # dominant kernel: the pair-energy kernel of the pipeline (two launches per
    # batch: first-bead sites, then orientation sites), timed with HIP events
    flops = widom_flops(r["screened_pairs"], r["evaluated_pairs"], M, r["nact"], r["n_ok"])
    eval_flops = r["evaluated_pairs"] * PAIR_FLOP
    ref_flops = reference_flops(n_fw, K, O, M - 1, M, r["nact"], r["n"], r["n_first_bead"], r["n_ok"])
    launches = max(1, r["launches"])
    avg_ms = r["kern_ms"] / launches
    achieved_tflops = eval_flops / launches / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else None
    batch_s = r["seconds"] / args.steps
    traffic, prof = load_traffic("void graspa::widom_eval_kernel<true")
    line = {
        "metric": "Widom insertions/sec on 1 GPU (CBMC 10x10, CO2, 1536-atom framework); GCMC MC moves/sec reported alongside",
        "value": value,
        "unit": "insertions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rv["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 pair terms (f32-seeded reciprocal, tabulated erfc(ar)/r), f32 overlap screen, f64 Ewald",
        "data": "synthetic (seeded jittered-lattice framework, Widom test insertions)",
        "config": {"workload": "Widom batch: 100,000 CO2 test insertions x (10 trial positions, 10 orientations), "
                               "1,536-atom framework, 24 A cubic, alpha 0.265, kmax 7, empty box",
                   "insertions_per_step": args.insertions, "framework_atoms": n_fw,
                   "first_bead_placed_fraction": r["n_first_bead"] / max(1, r["n"]),
                   "complete_fraction": r["n_ok"] / max(1, r["n"]),
                   "mean_rosenbluth_weight": r["sum_w"] / max(1, r["n"])},
        "roofline": {"bound": "valu",
                     "roof": "fp64 vector (VALU) peak: the pair terms compute in FP64 (the kernel uses no MFMA; "
                             "FP64 FMA issues at the unpacked FP32 rate on gfx950)",
                     "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": (achieved_tflops / FP64_PEAK_TFLOPS) if achieved_tflops else None,
                     "frac_of_fp32_peak": (achieved_tflops / FP32_PEAK_TFLOPS) if achieved_tflops else None,
                     "traffic": traffic,
                     "traffic_source": prof.get("source") if prof else None,
                     "kernel": "widom_eval_kernel<mixed> (pair energies of the surviving trial sites)",
                     "avg_launch_ms": avg_ms, "launches_per_batch": launches / args.steps,
                     "flop_per_launch": eval_flops / launches,
                     "evaluated_pairs_per_launch": r["evaluated_pairs"] / launches,
                     "batch_flop": flops / args.steps,
                     "screened_pairs_per_batch": r["screened_pairs"] / args.steps,
                     "batch_tflops": flops / args.steps / batch_s / 1e12,
                     "reference_algorithm_flop_per_batch": ref_flops / args.steps,
                     "reference_algorithm_equivalent_tflops": ref_flops / args.steps / batch_s / 1e12},
        "autotune": r["autotune"],
        "setup": dict(rv["setup"], note="per framework, outside the timed region: box upload with the screen's "
                                       "cell table, S(k) build, the first batch (host-built voxel maps), the "
                                       "autotuner"),
    }
    rho = s1.framework_density()
    from graspa_amd.deck import GAS_CONSTANT
    w = r["sum_w"] / max(1, r["n"])
    line["config"]["henry_mol_per_kg_pa"] = w / (GAS_CONSTANT * s1.temperature * rho)
    if args.batch_boxes > 0:
        line["gcmc_batched"] = run_batch_gcmc(args.batch_boxes, args.batch_cycles, engine.GRASPA_MIXED,
                                              warm_cycles=args.batch_warm_cycles)
    if not args.quick:
        # the same queued steps replayed from a captured HIP graph
        line["widom_insertions_per_s_queued"] = n_total / r["seconds"]
        line["widom_insertions_per_s_graph_replays"] = n_total / rv["seconds"] if use_graph else None
        line["widom_graph_same_totals_as_queued"] = (rv["sum_w"] == r["sum_w"] and rv["n_ok"] == r["n_ok"]
                                                     and rv["evaluated_pairs"] == r["evaluated_pairs"])
        rs = run_widom(s1, engine.GRASPA_MIXED, args.insertions, args.steps, 1, timing=False, queued=args.sync_steps)
        line["widom_insertions_per_s_" + ("queued" if args.sync_steps else "synchronous_steps")] = \
            args.insertions * args.steps / rs["seconds"]
        r64 = run_widom(s1, engine.GRASPA_FP64, args.insertions, max(1, args.steps // 5), 1, timing=False,
                        queued=not args.sync_steps)
        line["fp64_widom_insertions_per_s"] = args.insertions * max(1, args.steps // 5) / r64["seconds"]
        s_lo = synth.config1(n_fw=750)
        rl = run_widom(s_lo, engine.GRASPA_MIXED, args.insertions, max(1, args.steps // 2), 1,
                       queued=not args.sync_steps)
        fl = rl["evaluated_pairs"] * PAIR_FLOP
        al = rl["kern_ms"] / max(1, rl["launches"])
        line["low_density_variant"] = {
            "framework_atoms": 750, "insertions_per_s": args.insertions * max(1, args.steps // 2) / rl["seconds"],
            "complete_fraction": rl["n_ok"] / max(1, rl["n"]),
            "achieved_tflops": fl / max(1, rl["launches"]) / (al * 1e-3) / 1e12 if al > 0 else None}
        line["mixed_vs_fp64"] = mixed_error(s_lo, 4000)
        s0 = synth.config0(cycles=args.gcmc_cycles)
        g = run_gcmc(s0, args.gcmc_cycles)
        line["gcmc_fp64_single_box"] = g
        line["gcmc_device_driver_config0_fp64"] = run_batch_single(synth.config0(cycles=args.gcmc_cycles),
                                                                   args.gcmc_cycles, engine.GRASPA_FP64)
        line["gcmc_config2_butane_mixed"] = run_batch_single(synth.config2(), 200, engine.GRASPA_MIXED,
                                                             warm_cycles=100)
        # served stream: 500 us time slices per launch (each box at its own
        # pace), and the same jobs with a fixed 256 moves per box per launch
        line["served_stream"] = run_served(48, 20.0, 64, engine.GRASPA_MIXED, time_slice_us=500.0,
                                           moves_per_launch=4096)
        line["served_stream_fixed_256_moves"] = run_served(48, 20.0, 64, engine.GRASPA_MIXED)
        # GCMC where the moves do work: the reference's CO2-MFI deck, and the
        # 256-box distribution run for its 1,000 cycles to steady loading
        line["gcmc_co2_mfi"] = run_co2_mfi()
        line["gcmc_batched_steady"] = run_batch_steady(args.batch_boxes, 1000, engine.GRASPA_MIXED)
        # fixed dE error of the fast form in the batched configurations: the
        # same boxes and draws in both forms (24 boxes of the config-3
        # distribution, which the served stream also draws from, and the
        # butane box)
        line["mixed_vs_fp64_gcmc_config3"] = gcmc_de_error([synth.config3_box(b) for b in range(24)], 60,
                                                           [3000 + b for b in range(24)])
        line["mixed_vs_fp64_gcmc_config2"] = gcmc_de_error([synth.config2()], 300, [11], trace_cap=20000)
        # the 3,000-atom butane framework leaves no room for a molecule (every
        # trial overlaps), so the error is also measured on a 600-atom variant
        # of it, where insertions and moves happen
        line["mixed_vs_fp64_gcmc_config2_600_atom_variant"] = gcmc_de_error(
            [synth.config2(n_fw=600, P=1e5)], 300, [11], trace_cap=20000)
        # the reference's form of the same Widom workload: one insertion at a
        # time, host-driven, plain FP64 (one insertion per cycle, as the
        # Henry-coefficient decks run it)
        wp = run_gcmc(synth.config1(), 2000)
        line["widom_plain_single_box_fp64"] = dict(insertions=wp["moves"], seconds=wp["seconds"],
                                                   insertions_per_s=wp["moves_per_s"],
                                                   cycle_p50_ms=wp["cycle_p50_ms"], cycle_p99_ms=wp["cycle_p99_ms"])
        line["widom_batched_over_plain_fp64"] = value / wp["moves_per_s"]
        if "gcmc_batched" in line:
            line["gcmc_batched"]["over_single_box_fp64"] = line["gcmc_batched"]["moves_per_s"] / g["moves_per_s"]
        # at loading: the steady 256-box run over one CO2-MFI box (plain FP64,
        # device driver), both with molecules in the box
        st = line["gcmc_batched_steady"]
        st["over_single_box_fp64_co2_mfi"] = st["steady_moves_per_s"] / line["gcmc_co2_mfi"]["fp64_one_box"]["moves_per_s"]
        line["cpu_baseline"] = cpu_baseline(s1, args.cpu_sample)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
