#!/usr/bin/env python3
"""Throughput of the explicit MLS-MPM substep on one MI355X.

Workload: the Drucker-Prager sand column of 128 x 256 x 64 cells at dx = 1/512,
8 particles per cell (16,777,216 particles), seeded jitter, E = 3.5e5, nu = 0.3,
friction 30 deg, dt = 5e-5, sticky box walls, gravity -9.8 along y. One step is
one substep (table clear, activation, enlargement, bucket sort, grid clear,
P2G, grid update, G2P) on state resident in HBM, through
zsmpm_substep_sorted__rocm: G2P also writes the particles back in bucket order
(zpc's Vector::reorder), so every substep reads them block by block
(--in-place: zsmpm_substep__rocm, particles never move in memory).

The timed substeps are replays of a captured two-substep hipGraph
(ZSMPM_SUBSTEPS_GRAPH; --eager times the launches themselves). Before them, a
profiled eager pass of the same number of substeps records HIP events around
every stage on the stream it runs on; the per-stage times and the roofline of
the dominant kernel (algorithmic bytes per launch over its mean duration) come
from that pass.

Prints one JSON line: whole-job particle-substeps/s, the dominant kernel's
roofline, the same for the whole substep, and the host-core baseline (the
C++/OpenMP restatement of the reference CPU path, timed on a bounded sample of
the same workload).

    python bench.py [--steps K] [--warmup W] [--dims 128,256,64] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published

METRIC = "particle-substeps/s @16M particles + % HBM peak; surrogate vs plain"


def stage_bytes(stage: str, n: int, blocks: int, plastic: bool, table_size: int) -> float:
    """Algorithmic HBM bytes of one launch of a stage: every byte the algorithm
    must move once (particle state, permutation, grid channels), nothing for
    re-reads, hash probes or LDS."""
    nodes = 64 * blocks
    lj = 4 if plastic else 0
    if stage == "p2g":
        # read x, v, C, vol P F^T, m and the bucket permutation; flush m and the
        # force-folded momentum (4 channels) once per node
        return n * (12 + 12 + 36 + 36 + 4 + 4) + nodes * 16
    if stage == "g2p":
        # read x, F and the permutation; write x, v, C, F; gather v once per node
        return n * (12 + 36 + 4 + 12 + 12 + 36 + 36) + nodes * 12
    if stage == "stress":
        return n * (36 + lj + 36 + lj)  # read F (+ logJp), write vol P F^T (+ logJp)
    if stage == "grid_update":
        return nodes * (16 + 12)  # read m + momentum, write velocity
    if stage == "grid_clear":
        return nodes * 28
    if stage == "activate":
        return n * 12
    if stage == "bucket_sort":
        return n * (12 + 4)  # read x, write the permutation
    if stage == "table_clear":
        return table_size * 20
    if stage == "enlarge":
        return blocks * 12 * 8
    return 0.0


def substep_bytes(n: int, blocks: int, plastic: bool) -> float:
    """B_alg per substep: particle state read + written once (196 B, or 204 B
    with logJp), 16 B of sort traffic, 108 B per active grid node."""
    state = 204 if plastic else 196
    return n * (state + 16) + 108 * 64 * blocks


def measured_traffic(kernel_stage: str):
    """HBM bytes per launch of a stage's main kernel from a committed counter
    summary (profiles/traffic_<stage>_sand16M.txt: a line
    `bytes_per_launch <value>`), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"traffic_{kernel_stage}_sand16M.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and parts[0] == "bytes_per_launch":
            return float(parts[1])
    return None


def cpu_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n - 1)  # hardware_concurrency - 1, zpc's OpenMP default


def cpu_baseline(sc, warmup: int, steps: int) -> dict:
    """Time the OpenMP restatement on the host cores (a bounded sample)."""
    from oracle import oracle as O

    ps = O.HostParticles(sc.x, sc.v, sc.C, sc.F, sc.m, sc.logJp)
    threads = cpu_threads()
    secs, used, _ = O.omp_run(O.make_params(sc), ps, warmup, steps, threads)
    return {
        "value": sc.n * steps / secs,
        "unit": "particle-substeps/s",
        "cores": used,
        "kind": "port",
        "sample": f"{steps} timed substeps after {warmup} warm-up of the same {sc.n}-particle column "
                  f"(C++/OpenMP restatement of the reference CPU path, {used} threads, {secs:.1f} s)",
    }


def copy_peak(torch, nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Measured device-to-device copy bandwidth (bytes read + written per
    second), the practical HBM ceiling beside the 8 TB/s nominal peak."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    sec = a.elapsed_time(b) * 1e-3 / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / sec


def x_by_id(torch, sim):
    """Positions in original particle order (ids carried by the sorted driver)."""
    x = sim.parts.x
    if sim.parts.id is None:
        return x.clone()
    out = torch.empty_like(x)
    out[sim.parts.id.long()] = x
    return out


def positions_by_id(torch, rocm, sc, args, closure, substeps, widths, tiles):
    """Positions after `substeps` substeps of a second, untimed simulator that
    carries particle ids (the timed one moves zpc's attributes only: m, x, v,
    C, F, logJp), at the timed run's launch configuration; ids change nothing
    else, so this is the timed run's trajectory, found again by particle."""
    sim = rocm.Simulator(sc, reorder=not args.in_place, ids=True)
    if closure is not None:
        sim.pol.set_closure(closure[0], tol=closure[1], always_network=closure[2])
    sim.pol.set_launch_widths(*widths)
    sim.pol.set_launch_tiles(*tiles)
    sim.substeps(substeps, graph=not args.in_place)
    sim.sync()
    x = x_by_id(torch, sim)
    del sim
    return x


def run_path(torch, rocm, sc, args, multi, graph, closure=None, want_x=True):
    """Warm-up, a profiled eager pass and (graph) the timed replays of one
    simulator; closure: (params or None) selects the learned closure; want_x:
    also the positions by particle id (for the comparisons between legs)."""
    sim = rocm.Simulator(sc, reorder=not args.in_place, ids=False)
    if closure is not None:
        sim.pol.set_closure(closure[0], tol=closure[1], always_network=closure[2])
    tune = 0
    if multi:
        # 8 untimed substeps pick the P2G / G2P workgroup counts, whatever
        # --warmup is; then the warm-up substeps
        tune = 8
        if args.workgroups:
            sim.pol.set_launch_widths(*args.workgroups)
        sim.substeps(tune, autotune=not args.workgroups)
        if args.warmup > 0:
            sim.substeps(args.warmup)
    else:
        for _ in range(args.warmup):
            sim.substep()
    sim.sync()
    b0 = sim.nblocks()
    print(f"# warm-up done ({tune} tuning + {args.warmup} substeps), {b0} active blocks", file=sys.stderr, flush=True)
    ctr0 = sim.pol.counters()
    fb0, nw0 = int(ctr0[4]), int(ctr0[9])
    # profiled eager pass: every stage bracketed by HIP events on the stream it
    # runs on (per-stage and per-kernel times, the roofline)
    sim.pol.stage_times(reset=True)
    sim.pol.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if multi:
        sim.substeps(args.steps)
    else:
        for _ in range(args.steps):
            sim.substep()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sim.sync()
    sim.pol.profile(False)
    stages = sim.pol.stage_times(reset=True)
    eager_ms = (t1 - t0) * 1e3 / args.steps
    print(f"# profiled eager pass: {eager_ms:.3f} ms/substep", file=sys.stderr, flush=True)
    wall = t1 - t0
    if graph:
        # hipGraph replays of the captured two-substep pair; the capture and one
        # replay happen before the timed region
        sim.substeps(2, graph=True)
        sim.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.substeps(args.steps, graph=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sim.sync()
        wall = t1 - t0
        print(f"# graph replays: {wall * 1e3 / args.steps:.3f} ms/substep", file=sys.stderr, flush=True)
    b1 = sim.nblocks()
    substeps_run = tune + args.warmup + args.steps + (2 + args.steps if graph else 0)
    ctr = sim.pol.counters()
    widths, tiles = sim.pol.launch_widths(), sim.pol.launch_tiles()
    out = {"stages": stages, "blocks": 0.5 * (b0 + b1), "ms": wall * 1e3 / args.steps, "eager_ms": eager_ms,
           "launch_widths": widths, "launch_tiles": tiles,
           "fallbacks": int(ctr[4]) - fb0, "net_waves": int(ctr[9]) - nw0, "substeps": substeps_run}
    del sim
    if want_x:
        out["x"] = positions_by_id(torch, rocm, sc, args, closure, substeps_run, widths, tiles)
    return out


def _states(torch, rocm, scene, substeps):
    """(F, logJp) device tensors of a scene after `substeps` plain sorted substeps."""
    sim = rocm.Simulator(scene, reorder=True)
    if substeps > 0:
        sim.substeps(substeps)
    sim.sync()
    F = sim.parts.F.reshape(-1, 9).clone()
    lj = sim.parts.logJp.clone()
    del sim
    return F, lj


def _closure_stats(rocm, r, sc, n_pass):
    waves = n_pass * ((sc.n + 63) // 64)
    return {"network_wave_fraction": r["net_waves"] / waves,
            "fallback_fraction": r["fallbacks"] / (sc.n * n_pass)}


def surrogate_leg(torch, rocm, sc, args, graph, plain):
    """The same runs with the learned closure. The network is trained on the
    GPU (outside the timed regions) on a pool of about 50M states harvested
    from plain runs: the column after the warm-up, the column under a smooth
    initial strain (scenes.strained) at rest and after 20 substeps, plus
    random large-strain states. Legs on the column (where one Newton step from
    X0 = 0 certifies every wave, so the network is skipped) and on the
    strained column (5% compression and 3% shear, held by sticky walls on all
    six faces, where X0 = 0 certifies no particle): positions are compared with the plain
    run particle by particle after the same substeps, and the fraction of
    waves that evaluated the network and of particles that fell back to the
    SVD is reported."""
    from zpc_amd import closure as CL
    from zpc_amd import scenes as S

    t0 = time.perf_counter()
    strained = S.strained(sc, seed=5)
    parts = [_states(torch, rocm, sc, args.warmup), _states(torch, rocm, strained, 0),
             _states(torch, rocm, strained, 20)]
    Fs, ljs = CL.synthetic_states(1 << 18, 0)
    F = torch.cat([p[0] for p in parts] + [torch.as_tensor(Fs).cuda()]).contiguous()
    lj = torch.cat([p[1] for p in parts] + [torch.as_tensor(ljs).cuda()]).contiguous()
    del parts
    model = rocm.model_desc(sc.material)
    pol = rocm.Policy()
    tr = CL.train(pol, model, F, lj, steps=args.train_steps, batch=8192, lr=3e-3, seed=0)
    pool = int(F.shape[0])
    del F, lj
    torch.cuda.empty_cache()
    prep = time.perf_counter() - t0
    print(f"# closure trained on {pool} states: {args.train_steps} Adam steps in {tr.seconds:.2f} s, batch loss "
          f"{tr.losses[0]:.3e} -> {tr.losses[-1]:.3e}, kept step {tr.best_step} (mean loss {tr.eval_loss:.3e} on "
          f"{tr.eval_samples} states)", file=sys.stderr, flush=True)

    def leg(scene, base, params, always):
        r = run_path(torch, rocm, scene, args, True, graph, closure=(params, 1e-4, always))
        dx_err = (r["x"] - base["x"]).abs().max().item() / scene.dx
        n_pass = r["substeps"] - args.warmup - 8
        per_launch = {k: (ms / c if c else 0.0) for k, (ms, c) in r["stages"].items()}
        out = {
            "value": scene.n / (r["ms"] * 1e-3),
            "ms_per_step": r["ms"],
            "eager_ms_per_step": r["eager_ms"],
            "speedup_vs_plain": base["ms"] / r["ms"],
            "stress_ms_per_launch": per_launch.get("stress", 0.0),
            "plain_stress_ms_per_launch": base["stages"]["stress"][0] / max(base["stages"]["stress"][1], 1),
            "max_position_deviation_from_plain_dx": dx_err,
            "substeps_compared": r["substeps"],
        }
        out.update(_closure_stats(rocm, r, scene, n_pass))
        del r
        return out

    res = {}
    for name, params, always in (("trained", tr.params, False), ("trained_network_every_wave", tr.params, True),
                                 ("zero_start", None, False)):
        res[name] = leg(sc, plain, params, always)
        print(f"# surrogate ({name}): {res[name]['ms_per_step']:.3f} ms/substep, network waves "
              f"{res[name]['network_wave_fraction']:.3f}, fallback {res[name]['fallback_fraction']:.2e}, "
              f"max |x - x_plain| = {res[name]['max_position_deviation_from_plain_dx']:.2e} dx",
              file=sys.stderr, flush=True)
    # the strained column: the network is needed in most waves
    base = run_path(torch, rocm, strained, args, True, graph)
    print(f"# strained column, plain: {base['ms']:.3f} ms/substep", file=sys.stderr, flush=True)
    st = {"workload": "the same column at rest under a smooth initial strain, F = 0.95 I + 0.03 e_x e_y^T "
                      "+ U(-0.002, 0.002) per entry (seeded), in a closed box of sticky walls on its faces",
          "plain_ms_per_step": base["ms"], "plain_value": sc.n / (base["ms"] * 1e-3)}
    for name, params in (("trained", tr.params), ("zero_start", None)):
        st[name] = leg(strained, base, params, False)
        print(f"# strained column, surrogate ({name}): {st[name]['ms_per_step']:.3f} ms/substep, network waves "
              f"{st[name]['network_wave_fraction']:.3f}, fallback {st[name]['fallback_fraction']:.2e}, "
              f"max |x - x_plain| = {st[name]['max_position_deviation_from_plain_dx']:.2e} dx",
              file=sys.stderr, flush=True)
    del base
    res["strained_column"] = st
    res["network"] = ("MLP 9-64-64-7 GELU, BF16 MFMA / FP32 accumulate, trained on this GPU; 'trained' evaluates "
                      "it only in waves that one Newton step from X0 = 0 cannot certify a priori, "
                      "'trained_network_every_wave' in every wave, 'zero_start' never (uncertified particles "
                      "take the SVD fallback)")
    res["training"] = {"adam_steps": args.train_steps, "batch": 8192, "samples": tr.samples, "pool": pool,
                       "pool_sources": "column after the warm-up, strained column at t = 0 and after 20 plain "
                                       "substeps (16,777,216 states each), 262,144 random states",
                       "seconds": tr.seconds, "loss_first": tr.losses[0], "loss_last": tr.losses[-1],
                       "loss_eval": tr.eval_loss, "kept_step": tr.best_step, "eval_samples": tr.eval_samples,
                       "data_seconds": prep - tr.seconds}
    return res


def batched_leg(torch, rocm, args) -> dict:
    """The 64-scene workload: independent scenes of 10K..500K particles
    (boxes and spheres, fixed-corotated and sand, dx 1/128 and 1/256) in one
    table and one grid, advanced by zsmpm_substeps_batched__rocm with every
    stage one launch over all scenes: 8 autotuning substeps and 2 warm-up
    substeps, then the timed substeps as replays of the captured two-substep
    graph. With --batch-per-scene the same scenes are also timed the way a
    single-scene driver would run them (each its own policy and captured
    graph, one scene after another) for comparison."""
    from zpc_amd import scenes as S

    scenes = S.batch_scenes(64)
    n = sum(sc.n for sc in scenes)
    b = rocm.Batch(scenes)
    b.substeps(8, autotune=True)
    b.substeps(2, graph=True)
    b.sync()
    blocks = b.nblocks()
    steps = args.batch_steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.substeps(steps, graph=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    b.sync()
    ms = (t1 - t0) * 1e3 / steps
    widths = b.pol.launch_widths()
    tiles = b.pol.launch_tiles()
    print(f"# 64 scenes, {n} particles, {blocks} blocks: {ms:.3f} ms per substep of every scene "
          f"(one launch per stage)", file=sys.stderr, flush=True)
    largest = max(scenes, key=lambda sc: sc.n)
    out = {
        "workload": "64 scenes, particle counts log-uniform in [10K, 500K] (seeded), jittered boxes and spheres, "
                    "fixed-corotated / Drucker-Prager, dx 1/128 or 1/256, dt 1e-4",
        "particles": n,
        "largest_scene_particles": largest.n,
        "active_blocks": blocks,
        "value": n / (ms * 1e-3),
        "unit": "particle-substeps/s",
        "ms_per_step": ms,
        "steps": steps,
        "p2g_g2p_workgroups": list(widths),
        "p2g_chunk_g2p_threads": list(tiles),
        "schedule": "zsmpm_substeps_batched__rocm: one table and grid for all scenes, every stage one launch, "
                    "replays of a captured two-substep hipGraph",
    }
    del b
    torch.cuda.empty_cache()
    if args.batch_per_scene:
        sims = [rocm.Simulator(sc, reorder=True) for sc in scenes]
        for sim in sims:
            sim.substeps(8, autotune=True)
            sim.substeps(2, graph=True)
        for sim in sims:
            sim.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for sim in sims:
            sim.substeps(steps, graph=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms1 = (t1 - t0) * 1e3 / steps
        print(f"# per-scene graphs: {ms1:.3f} ms per substep of every scene", file=sys.stderr, flush=True)
        out["per_scene_graphs"] = {"ms_per_step": ms1, "value": n / (ms1 * 1e-3),
                                   "schedule": "each scene its own policy and captured graph, one after another"}
        out["speedup_vs_per_scene"] = ms1 / ms
        del sims
        torch.cuda.empty_cache()
    return out


def adapters_leg(torch, rocm, sc, reps: int = 20):
    """The layout adapters at the bench scene's size: the particles into and
    out of a TileVector<f32, 32> with zpc's particle tags (m, x, v, F, C,
    logJp, id; 27 channels, F and C row-major), and a random permutation of
    its tiles (TileVector::reorderTiles). Each is timed with HIP events on the
    policy's stream over `reps` launches; bytes are every attribute read once
    and written once (108 B per particle each way, 2 x 108 B per element for
    the tile reorder)."""
    pol = rocm.Policy()
    parts = rocm.DeviceParticles.from_scene(sc, ids=True)
    tags = [("m", 1), ("x", 3), ("v", 3), ("F", 9), ("C", 9), ("logJp", 1), ("id", 1)]
    tiles = rocm.DeviceTiles(tags, sc.n, 32)
    other = rocm.DeviceTiles(tags, sc.n, 32)
    back = parts.empty_like()
    perm = torch.randperm(tiles.ntiles, device="cuda", dtype=torch.int64).to(torch.int32)
    per_particle = 4 * tiles.nch

    def timed(fn):
        fn()
        pol.sync()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(pol.stream)
        for _ in range(reps):
            fn()
        b.record(pol.stream)
        b.synchronize()
        pol.sync()
        return a.elapsed_time(b) / reps

    out = {"workload": f"{sc.n} particles of the bench scene in a TileVector<f32, 32> of 27 channels "
                       "(m, x, v, F, C, logJp, id)"}
    for name, fn, nbytes in (
            ("to_tiles", lambda: pol.particles_to_tiles(parts, tiles), 2 * per_particle * sc.n),
            ("from_tiles", lambda: pol.particles_from_tiles(tiles, back), 2 * per_particle * sc.n),
            ("tiles_reorder", lambda: pol.tiles_reorder(tiles, other, perm), 2 * 4 * tiles.data.numel())):
        ms = timed(fn)
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "bytes": nbytes, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
        print(f"# {name}: {ms:.3f} ms, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    h0, h1 = parts.host(), back.host()
    out["round_trip_exact"] = all(bool((h0[k] == h1[k]).all()) for k in ("x", "v", "C", "F", "m", "logJp", "id"))
    del parts, tiles, other, back
    pol.close()
    torch.cuda.empty_cache()
    return out


def sparse_grid_leg(torch, rocm, sc, reps: int = 10):
    """The SparseGrid on zpc's bucketed hash table at the bench scene's size:
    - activation of SparseGrid<3, f32, 8> blocks from all particle positions
      (one key per particle, a few thousand distinct blocks);
    - insertion of 16M random keys with 4M distinct ones into a fresh table,
      and 16M queries of present keys (16-lane bucket compares);
    - trilinear samples of a signed-distance channel at every particle;
    - the grid update of the bench scene with and without a level-set
      boundary (a planar signed distance under the column: slip floor).
    Times are HIP events on the policy's stream (insertions synchronise once
    per placement round, so their wall time is what a caller sees)."""
    import numpy as np

    pol = rocm.Policy()
    out = {"workload": f"{sc.n} particles of the bench scene; 2^24 random keys with 2^22 distinct"}
    x = torch.as_tensor(sc.x).cuda()
    dx8 = sc.dx  # one SparseGrid voxel per MPM cell
    i2w = np.eye(4)
    i2w[:3, :3] *= dx8

    def ev_time(fn, n=reps):
        fn()
        pol.sync()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(pol.stream)
        for _ in range(n):
            fn()
        b.record(pol.stream)
        b.synchronize()
        pol.sync()
        return a.elapsed_time(b) / n

    g = rocm.DeviceSparseGrid(8192, 1, i2w, background=1.0, pol=pol)

    def activate():
        g.table.clear()
        g.activate(x)

    ms = ev_time(activate)
    blocks = g.table.size()
    out["activate"] = {"ms": ms, "keys_per_s": sc.n / (ms * 1e-3), "blocks": blocks}
    print(f"# spg activate: {ms:.3f} ms for {sc.n} positions, {blocks} blocks", file=sys.stderr, flush=True)
    # a planar signed distance y - y_floor on every cell
    coords = g.cell_coords()
    y_floor = float(sc.x[:, 1].min()) + 0.5 * sc.dx
    vals = ((coords[..., 1].astype(np.float32) * np.float32(dx8)) - np.float32(y_floor)).reshape(blocks, 1, 512)
    g.data[:blocks].copy_(torch.as_tensor(vals).cuda())
    samp = torch.empty(sc.n, dtype=torch.float32, device="cuda")

    def sample():
        check_ = rocm.check
        check_(pol.L.zsmpm_spg_sample__rocm(pol.h, rocm.C.byref(g.desc()), 0, 1, x.data_ptr(), sc.n,
                                            samp.data_ptr()), "spg_sample")

    ms = ev_time(sample)
    out["sample"] = {"ms": ms, "samples_per_s": sc.n / (ms * 1e-3)}
    print(f"# spg sample: {ms:.3f} ms", file=sys.stderr, flush=True)

    n_keys, n_distinct = 1 << 24, 1 << 22
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    distinct = torch.randint(-(1 << 20), 1 << 20, (n_distinct, 3), dtype=torch.int32, device="cuda", generator=gen)
    keys = distinct[torch.randint(0, n_distinct, (n_keys,), device="cuda", generator=gen)].contiguous()
    t = rocm.DeviceBht(rocm.bht_table_size(n_distinct), pol)
    res = torch.empty(n_keys, dtype=torch.int32, device="cuda")

    def insert():
        t.clear()
        t.insert(keys, out=res)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    insert()
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(3):
        insert()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / 3
    pol.sync()
    rounds = int(pol.counters()[10])
    out["bht_insert"] = {"ms": ms, "keys_per_s": n_keys / (ms * 1e-3), "distinct": t.size(), "rounds": rounds,
                         "table_size": t.T, "first_call_ms": first * 1e3}
    print(f"# bht insert: {ms:.3f} ms for {n_keys} keys ({t.size()} distinct, {rounds} rounds)", file=sys.stderr,
          flush=True)
    q = keys

    def query():
        t.query(q, out=res)

    ms = ev_time(query)
    bytes_q = n_keys * (12 + 4 + 256)  # key, result, one bucket of 16 storage keys
    out["bht_query"] = {"ms": ms, "keys_per_s": n_keys / (ms * 1e-3), "GB/s_one_bucket": bytes_q / (ms * 1e-3) / 1e9}
    print(f"# bht query: {ms:.3f} ms", file=sys.stderr, flush=True)
    del keys, distinct, res, t

    sim = rocm.Simulator(sc, reorder=True)
    sim.substeps(2)
    sim.sync()

    def update():
        sim.grid_update(fold=0)

    ms0 = ev_time(update, 20)
    sim.pol.set_levelset_boundary(g, 0, -1, 1)
    ms1 = ev_time(update, 20)
    sim.pol.set_levelset_boundary(None)
    out["grid_update_levelset"] = {"ms_without": ms0, "ms_with": ms1,
                                   "boundary": "planar signed distance y - y_floor in a SparseGrid over the column, "
                                               "slip"}
    print(f"# grid update: {ms0:.3f} ms, with the level-set boundary {ms1:.3f} ms", file=sys.stderr, flush=True)
    del sim, g, x, samp
    pol.close()
    torch.cuda.empty_cache()
    return out


def implicit_leg(torch, rocm, sc, reps: int = 5):
    """The implicit step's operators on the bench scene after two sorted
    substeps (particles in bucket order):
    G2P2G of the grid velocity (a per-particle gather of the 27-node stencil,
    F update and stress, then P2G's LDS-binned scatter of the force), the
    system's multiply, and a
    conjugate-gradient solve of (f dt^2 + m) v = m v_explicit to zpc's
    tolerances (tol 1e-6, rel_tol 0.5)."""
    import numpy as np

    sim = rocm.Simulator(sc, reorder=True)  # particles in bucket order, as the sorted driver leaves them
    sim.substeps(2)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.clear_grid()
    sim.p2g()
    sim.sync()
    nb = sim.nblocks()
    n = nb * 64 * 3
    g = sim.grid.data[:nb]
    m = g[:, 0, :].reshape(-1).repeat_interleave(3)
    mv = g[:, 1:4, :].permute(0, 2, 1).reshape(-1)
    v = sim.dof_vector()
    v[:n] = torch.where(m > 0, mv / torch.where(m > 0, m, torch.ones_like(m)), torch.zeros_like(m))
    b = sim.dof_vector()
    b[:n] = m * v[:n]
    r = sim.dof_vector()

    def ev_time(fn):
        fn()
        sim.sync()
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(sim.pol.stream)
        for _ in range(reps):
            fn()
        e.record(sim.pol.stream)
        e.synchronize()
        sim.sync()
        return a.elapsed_time(e) / reps

    ms_g2p2g = ev_time(lambda: sim.g2p2g(v, r))
    ms_mult = ev_time(lambda: sim.implicit_multiply(v, r))
    x = sim.dof_vector()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it, res = sim.implicit_solve(x, b)
    torch.cuda.synchronize()
    ms_solve = (time.perf_counter() - t0) * 1e3
    out = {"workload": f"{sc.n} particles of the bench scene after two substeps, {nb} blocks",
           "g2p2g_ms": ms_g2p2g, "g2p2g_particles_per_s": sc.n / (ms_g2p2g * 1e-3), "multiply_ms": ms_mult,
           "solve_ms": ms_solve, "solve_iterations": it, "solve_residual": res}
    print(f"# implicit: g2p2g {ms_g2p2g:.3f} ms, multiply {ms_mult:.3f} ms, solve {ms_solve:.2f} ms "
          f"({it} iterations, residual {res:.3e})", file=sys.stderr, flush=True)
    del sim, v, b, r, x
    torch.cuda.empty_cache()
    return out


def transfers_leg(torch, rocm, sc, reps: int = 5):
    """The particle -> cell -> grid transfers beside the APIC ones on the bench
    scene after two sorted substeps (particles in bucket order): P2C2G (stress pass +
    cell-kernel scatter) against P2G (stress pass + quadratic scatter), and
    G2C2P against G2P, each a whole C-ABI call on the policy's stream."""
    sim = rocm.Simulator(sc, reorder=True)  # particles in bucket order, as the sorted driver leaves them
    sim.substeps(2)
    sim.clear_table()
    sim.activate()
    sim.enlarge()
    sim.bucket_sort()
    sim.sync()

    def ev_time(fn):
        fn()
        sim.sync()
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(sim.pol.stream)
        for _ in range(reps):
            fn()
        e.record(sim.pol.stream)
        e.synchronize()
        sim.sync()
        return a.elapsed_time(e) / reps

    def p2(fn):
        def run():
            sim.clear_grid()
            fn()
        return run

    ms_p2g = ev_time(p2(sim.p2g))
    ms_p2c2g = ev_time(p2(sim.p2c2g))
    sim.grid_update(fold=0)
    ms_g2p = ev_time(sim.g2p)
    ms_g2c2p = ev_time(sim.g2c2p)
    out = {"workload": f"{sc.n} particles of the bench scene after two substeps, {sim.nblocks()} blocks",
           "p2c2g_ms": ms_p2c2g, "p2g_ms": ms_p2g, "g2c2p_ms": ms_g2c2p, "g2p_ms": ms_g2p,
           "p2c2g_particles_per_s": sc.n / (ms_p2c2g * 1e-3), "g2c2p_particles_per_s": sc.n / (ms_g2c2p * 1e-3)}
    del sim
    torch.cuda.empty_cache()
    # whole sorted substeps (hipGraph replays) with each transfer scheme, from
    # the column at rest (its C = 0 is B = 0 for the cell scheme)
    for name, scheme in (("apic", rocm.TRANSFER_APIC), ("cell", rocm.TRANSFER_CELL)):
        sim = rocm.Simulator(sc, reorder=True)
        sim.pol.set_transfer_scheme(scheme)
        sim.substeps(8, autotune=True)
        sim.substeps(4, graph=True)
        sim.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.substeps(20, graph=True)
        torch.cuda.synchronize()
        out[f"substep_{name}_ms"] = (time.perf_counter() - t0) * 1e3 / 20
        del sim
        torch.cuda.empty_cache()
    print(f"# transfers: p2c2g {ms_p2c2g:.3f} ms (p2g {ms_p2g:.3f}), g2c2p {ms_g2c2p:.3f} ms (g2p {ms_g2p:.3f}); "
          f"sorted substep with the cell transfers {out['substep_cell_ms']:.3f} ms "
          f"(APIC {out['substep_apic_ms']:.3f})", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dims", default="128,256,64", help="sand column cells (x,y,z)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-core baseline")
    ap.add_argument("--in-place", action="store_true",
                    help="keep the particles in their original order (no reorder after G2P)")
    ap.add_argument("--single", action="store_true",
                    help="one zsmpm_substep_sorted__rocm call per substep instead of one multi-substep call")
    ap.add_argument("--eager", action="store_true",
                    help="time the profiled eager launches instead of hipGraph replays")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-closure leg")
    ap.add_argument("--train-steps", type=int, default=4000, help="Adam steps of the closure network")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched leg")
    ap.add_argument("--batch-steps", type=int, default=100, help="timed substeps of the batched leg")
    ap.add_argument("--batch-per-scene", action="store_true",
                    help="also time the 64 scenes as separate single-scene graphs, one after another")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many bytes (the reference kernel "
                         "that scales FETCH_SIZE / WRITE_SIZE counters to bytes)")
    ap.add_argument("--no-adapters", action="store_true", help="skip the layout adapters' leg")
    ap.add_argument("--no-sparse-grid", action="store_true", help="skip the SparseGrid / bht leg")
    ap.add_argument("--no-implicit", action="store_true", help="skip the implicit step's leg")
    ap.add_argument("--no-transfers", action="store_true", help="skip the P2C2G / G2C2P leg")
    ap.add_argument("--workgroups", default="",
                    help="P2G,G2P workgroup counts to launch with in place of the autotuned ones (the 8 "
                         "tuning substeps still run, untuned), e.g. to hold a counter pass to a timed run's widths")
    args = ap.parse_args()
    args.workgroups = tuple(int(v) for v in args.workgroups.split(",")) if args.workgroups else ()
    if args.workgroups and len(args.workgroups) != 2:
        ap.error("--workgroups takes two counts: P2G,G2P")

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the substep runs on a single MI355X here; see the sharding notes in the design document)",
              file=sys.stderr)
        return 2
    if args.steps <= 0 or args.warmup < 0:
        print("--steps must be positive and --warmup non-negative", file=sys.stderr)
        return 2

    import torch

    from zpc_amd import rocm
    from zpc_amd import scenes as S

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1

    dims = tuple(int(v) for v in args.dims.split(","))
    sc = S.sand_scene(dims=dims)
    plastic = sc.logJp is not None
    print(f"# scene {sc.name}: {sc.n} particles, table {sc.table_size}", file=sys.stderr, flush=True)

    if args.copy_bytes > 0:
        words = args.copy_bytes // 4
        src = torch.ones(words, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy kernel: {4 * words} bytes read, {4 * words} bytes written", file=sys.stderr, flush=True)
        del src, dst

    copy_gbs = copy_peak(torch) / 1e9
    print(f"# device copy: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    multi = not (args.in_place or args.single)
    graph = multi and not args.eager
    plain = run_path(torch, rocm, sc, args, multi, graph, want_x=not args.no_surrogate)
    stages, blocks, ms_per_step, eager_ms = plain["stages"], plain["blocks"], plain["ms"], plain["eager_ms"]
    value = sc.n / (ms_per_step * 1e-3)
    surrogate = None
    if not args.no_surrogate and multi:
        surrogate = surrogate_leg(torch, rocm, sc, args, graph, plain)
    x_plain = plain.pop("x", None)
    batched = None
    if not args.no_batch:
        del x_plain
        x_plain = None
        torch.cuda.empty_cache()
        batched = batched_leg(torch, rocm, args)
    adapters = None
    if not args.no_adapters:
        if x_plain is not None:
            del x_plain
            x_plain = None
            torch.cuda.empty_cache()
        adapters = adapters_leg(torch, rocm, sc)
    sparse_grid = None
    if not args.no_sparse_grid:
        if x_plain is not None:
            del x_plain
            x_plain = None
            torch.cuda.empty_cache()
        sparse_grid = sparse_grid_leg(torch, rocm, sc)
    implicit = None
    if not args.no_implicit:
        if x_plain is not None:
            del x_plain
            x_plain = None
            torch.cuda.empty_cache()
        implicit = implicit_leg(torch, rocm, sc)
    transfers = None
    if not args.no_transfers:
        if x_plain is not None:
            del x_plain
            x_plain = None
            torch.cuda.empty_cache()
        transfers = transfers_leg(torch, rocm, sc)
    per_launch = {k: (ms / calls if calls else 0.0) for k, (ms, calls) in stages.items()}
    dom = max(per_launch, key=lambda k: stages[k][0])
    dom_bytes = stage_bytes(dom, sc.n, blocks, plastic, sc.table_size)
    dom_gbs = dom_bytes / (per_launch[dom] * 1e-3) / 1e9
    sub_bytes = substep_bytes(sc.n, blocks, plastic)
    sub_gbs = sub_bytes / (ms_per_step * 1e-3) / 1e9
    print("# stage ms/launch: " + ", ".join(f"{k} {v:.3f}" for k, v in per_launch.items()),
          file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded jittered sand column (splitmix64), initial state F=I, C=0, v=0",
        "config": {
            "workload": f"Drucker-Prager sand column {dims[0]}x{dims[1]}x{dims[2]} cells, 8 ppc, dx=1/512, "
                        f"{sc.n} particles, dt=5e-5, sticky walls, plain SVD return mapping",
            "particles": sc.n,
            "active_blocks": blocks,
            "table_size": sc.table_size,
            "path": "plain",
            "particle_order": "original (in place)" if args.in_place else
                              "bucket order, reordered inside G2P every substep",
            "particle_attributes": "m, x, v, C, F, logJp (zpc's; no particle id)",
            "p2g_g2p_workgroups": list(plain["launch_widths"]),
            "p2g_chunk_g2p_threads": list(plain["launch_tiles"]),
            "driver": ("zsmpm_substeps_sorted__rocm with ZSMPM_SUBSTEPS_GRAPH (a captured two-substep "
                       "hipGraph replayed)" if graph else
                       "zsmpm_substeps_sorted__rocm (one call for all timed substeps)") if multi else
                      ("zsmpm_substep__rocm per substep" if args.in_place else
                       "zsmpm_substep_sorted__rocm per substep"),
        },
        "eager_ms_per_step": eager_ms,
        "stage_timing": f"HIP events around every stage of a profiled eager pass of the same {args.steps} substeps",
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": dom_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom) if dims == (128, 256, 64) else None,
            "bytes_per_launch": dom_bytes,
            "ms_per_launch": per_launch[dom],
            "copy_peak": copy_gbs,
            "frac_of_copy_peak": dom_gbs / copy_gbs,
        },
        "substep_roofline": {
            "bytes_per_substep": sub_bytes,
            "achieved": sub_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": sub_gbs / HBM_PEAK_GBS,
            "copy_peak": copy_gbs,
            "frac_of_copy_peak": sub_gbs / copy_gbs,
        },
        "stage_ms_per_launch": per_launch,
        "surrogate": surrogate,
        "batched": batched,
        "adapters": adapters,
        "sparse_grid": sparse_grid,
        "implicit": implicit,
        "transfers": transfers,
        "cpu_baseline": None,
    }
    del x_plain
    torch.cuda.synchronize()
    if not args.no_cpu:
        print(f"# cpu baseline: {args.cpu_warmup} + {args.cpu_steps} substeps on {cpu_threads()} threads",
              file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_warmup, args.cpu_steps)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
