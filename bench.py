"""Benchmark of the MI355X RHS + RK3 path (one GPU).

Workload (default): single-level 256^3 FP64 reacting flow through a porous solid:
seeded packed spheres (radii U[6,14] cells, porosity 0.6) as the immersed boundary,
5 species (4 gas + 1 site), 2 Arrhenius reactions, inflow u = 1 m/s at T = 700 K,
WENO5 fluxes (order_hyp = 5), SSP-RK3, viscous/heat/species diffusion, explicit
chemistry, CFL 0.3. A "step" is one RK3 step of the whole level with inputs
resident in HBM.

Prints one JSON line: cell-updates/s (value), the roofline of the dominant kernel
measured with HIP events on the stream it runs on, and the CPU oracle timed on a
bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "FP64 cell-updates/s per RK3 step, % HBM roofline; adjoint/forward cost ratio"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(name):
    from mesoflow_amd import cases
    if name == "porous256":
        # 32^3 boxes: amr.max_grid_size of the reference's porous-particle model
        # (models/porousParticle/inputs:17)
        c = cases.porous_reactor(n_cell=(256, 256, 256), nspec=5, nreact=2, geometry="spheres", porosity=0.6,
                                 rmin=6.0, rmax=14.0, order_hyp=5, cfl=0.3, seed=20261203, max_grid_size=32)
        desc = ("porous reactor 256^3 FP64, 5 species, 2 Arrhenius reactions, IB sphere pack r~U[6,14] "
                "porosity 0.6, inflow u=1 m/s T=700 K, WENO5 (order_hyp=5), single level of 512 boxes of 32^3, SSP-RK3")
    elif name == "uneven256":
        boxes = cases.random_boxes((256, 256, 256), 8, 64, seed=7, p_stop=0.5)
        c = cases.porous_reactor(n_cell=(256, 256, 256), nspec=5, nreact=2, geometry="grf", porosity=0.5,
                                 corr_len=8.0, order_hyp=5, cfl=0.3, seed=20261205, boxes=boxes)
        desc = ("porous reactor 256^3 FP64, 5 species, 2 reactions, GRF IB porosity 0.5, WENO5, single level "
                "of %d uneven boxes with edges in [8,64], SSP-RK3" % len(boxes))
    elif name == "sod64":
        c = cases.shocktube_n2(n_cell=(64, 64, 64), prob_hi=(1.0, 1.0, 1.0), order_hyp=5, cfl=0.5)
        desc = "Sod shock tube 64^3 FP64, 1 species, WENO5, SSP-RK3 (CPU-runnable reference case)"
    elif name == "porous512":
        c = cases.porous_reactor(n_cell=(512, 512, 512), nspec=10, nreact=8, geometry="grf", porosity=0.45,
                                 corr_len=8.0, order_hyp=5, cfl=0.3, seed=20261207)
        desc = "porous reactor 512^3 FP64, 10 species, 8 Arrhenius reactions, GRF IB porosity 0.45, WENO5, SSP-RK3"
    else:
        raise SystemExit("unknown config %s" % name)
    return c, desc


def nv_of(case):
    return 12 + case.nspec


def host_cpu():
    """CPU model, logical CPUs of the host and those this process may use"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_baseline(name, budget_s):
    """Oracle (CPU restatement, OpenMP over every CPU this process may use,
    capped by OMP_NUM_THREADS when set) on a bounded sample of the workload."""
    from mesoflow_amd import cases
    from oracle.pyoracle import OracleLevel
    model, nproc, usable = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", usable))
    if name == "sod64":
        c = cases.shocktube_n2(n_cell=(64, 64, 64), prob_hi=(1.0, 1.0, 1.0), order_hyp=5, cfl=0.5)
        sample = "Sod 64^3, WENO5"
    else:
        n = 128
        c = cases.porous_reactor(n_cell=(n, n, n), nspec=5 if name != "porous512" else 10,
                                 nreact=2 if name != "porous512" else 8,
                                 geometry="spheres" if name != "porous512" else "grf", porosity=0.6 if name != "porous512" else 0.45,
                                 rmin=6.0, rmax=14.0, corr_len=8.0, order_hyp=5, cfl=0.3, seed=20261203)
        sample = "same physics on a 128^3 sub-problem (1/8 of the cells; sphere pack, NS=%d)" % c.nspec
    L = OracleLevel(c)
    L.set_state(c.make_state())
    cells = float(np.prod(c.n_cell))
    t0 = time.time()
    L.step(1)
    t1 = time.time()
    per = max(t1 - t0, 1e-3)
    nmore = max(1, min(50, int(budget_s / per) - 1))
    t2 = time.time()
    L.step(nmore)
    t3 = time.time()
    rate = cells * nmore / (t3 - t2)
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port", "cpu_model": model,
            "nproc": nproc, "cpus_usable": usable,
            "sample": "%s, %d RK3 steps after 1 warm-up step, %.1f s, OMP threads %d of %d usable (%d on the host)"
                      % (sample, nmore, t3 - t2, threads, usable, nproc)}


def adjoint_bench(nsteps=200, ncheck_binomial=20):
    """Discrete adjoint of the porous reactor at 128^3 (5 species, 3 reactions
    with fast kinetics, sphere pack, perturbed initial state): time of the
    plain forward per step, of mf_adjoint_forward and of the reverse sweep
    (mf_adjoint_backward) per step, with every state kept (the schedule
    degenerates to store-all when the snapshots fit in HBM) and with
    ncheck_binomial binomial checkpoints (forward segments recomputed)."""
    import torch
    from mesoflow_amd import abi, cases, runtime
    c = cases.porous_adjoint()
    s0 = cases.perturbed_state(c)
    ctx0 = runtime.Context(c)
    L0 = runtime.Level(ctx0)
    L0.set_state(s0)
    dt = L0.est_dt()
    L0.close()
    ctx0.close()
    c.params.fixed_dt = 1
    c.params.dt_fixed = dt
    ctx = runtime.Context(c)
    L = runtime.Level(ctx)
    L.set_state(s0)
    L.autotune()

    def fresh():
        L.set_state(s0)
        L.set_time(0.0, 1.0e100, 0)
        torch.cuda.synchronize()

    fresh()
    L.step(2, abi.MODE_BATCHED)
    fresh()
    t0 = time.perf_counter()
    L.step(nsteps, abi.MODE_BATCHED)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / nsteps
    out = {"workload": "porous reactor 128^3 FP64, 5 species, 3 Arrhenius reactions (A 1e12..3.7e11, Ea 60-64 kJ/mol), "
                       "IB sphere pack r~U[6,14] porosity 0.6, perturbed initial state, WENO5, fixed dt, %d RK3 steps; "
                       "J = time-integrated outlet mass flux of species 1; gradient w.r.t. the 3 log pre-exponentials "
                       "and the inflow velocity" % nsteps,
           "forward_ms_per_step": fwd * 1e3}
    padded = (128 + 6) ** 3 * c.nv * 8
    free, _ = torch.cuda.mem_get_info()
    runs = []
    if nsteps * padded < 0.6 * free:
        runs.append(("store_all", nsteps))
    runs.append(("binomial", ncheck_binomial))
    for name, ncheck in runs:
        fresh()
        t0 = time.perf_counter()
        J = L.adjoint_forward(nsteps, ncheck, 1)
        t1 = time.perf_counter()
        gA, gv = L.adjoint_backward()
        t2 = time.perf_counter()
        bwd = (t2 - t1) / nsteps
        out[name] = {"ncheck": ncheck, "adjoint_forward_ms_per_step": (t1 - t0) / nsteps * 1e3,
                     "backward_ms_per_step": bwd * 1e3, "backward_over_forward": bwd / fwd,
                     "J": J, "J_unit": "kg", "dJ_dlogA": [float(x) for x in gA], "dJ_dinvel": gv,
                     "dlnJ_dlogA": [float(x) / J if J != 0.0 else None for x in gA]}
        log("adjoint %s: %s" % (name, json.dumps(out[name])))
    best = min(out[n]["backward_over_forward"] for n, _ in runs)
    out["adjoint_over_forward"] = best
    L.close()
    ctx.close()
    return out


def substeps(dt_react, dt_rk):
    """Substeps of the reference's `while (tchem < dt_react) {...; tchem += dt_rk;}`."""
    n, t = 0, 0.0
    while t < dt_react:
        n += 1
        t += dt_rk
    return n


def split_chem_cpu(name, nsub=100, dt_rk=1.0e-6):
    """Chemistry phase of the split mode on the oracle (64^3 sample of the workload)."""
    from mesoflow_amd import cases
    from oracle.pyoracle import OracleLevel
    c = cases.porous_reactor(n_cell=(64, 64, 64), nspec=5 if name != "porous512" else 10,
                             nreact=2 if name != "porous512" else 8, geometry="spheres", porosity=0.6, rmin=6.0, rmax=14.0,
                             order_hyp=5, cfl=0.3, seed=20261203)
    L = OracleLevel(c)
    L.set_state(c.make_state())
    t0 = time.time()
    L.advance_chemistry(0.0, nsub * dt_rk, dt_rk)
    secs = time.time() - t0
    return 64 ** 3 * substeps(nsub * dt_rk, dt_rk) / secs


def read_traffic(config):
    """(HBM bytes per face-sweep launch, the profile's tile line) from the
    committed rocprofv3 --pmc summary of this workload"""
    path = os.path.join(ROOT, "profiles", "traffic_%s.txt" % config)
    if not os.path.exists(path):
        return None, None
    val, tile = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            val = float(line.split(":")[1])
        if line.startswith("tile:"):
            tile = line.split(":", 1)[1].strip()
    return val, tile


def box_rhs_bench(ctx, lvl, reps, edge=32):
    """The drop-in per-box call (mf_box_rhs with face-flux output) on one
    edge^3 box of the level's ghost-filled state: ms per call and cell-RHS
    evaluations per second"""
    import torch
    from mesoflow_amd import abi, runtime
    nv = ctx.nv
    pv = abi.Array4()
    runtime.check(runtime.lib().mf_level_state_view(lvl.h, pv), "mf_level_state_view")
    lvl.fill_ghosts(0.0)
    ctx.sync()
    lo = (0, 0, 0)
    hi = (edge - 1,) * 3
    ds = torch.zeros((nv, edge, edge, edge), dtype=torch.float64, device="cuda")
    fl, fv = [], (abi.Array4 * 3)()
    for d in range(3):
        m = [edge] * 3
        m[d] += 1
        t = torch.zeros((nv, m[2], m[1], m[0]), dtype=torch.float64, device="cuda")
        fl.append(t)
        fv[d] = abi.make_array4(t.data_ptr(), lo, [lo[q] + m[q] for q in range(3)], nv)
    dsv = abi.make_array4(ds.data_ptr(), lo, [edge] * 3, nv)
    torch.cuda.synchronize()

    def call():
        runtime.check(runtime.lib().mf_box_rhs(ctx.h, abi.make_box(lo, hi), pv, dsv, fv, 0.0), "mf_box_rhs")

    call()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    ctx.sync()
    ms = (time.perf_counter() - t0) / reps * 1e3
    return {"box": "%d^3 of the workload's state, flux views given, all components" % edge, "reps": reps,
            "ms_per_call": round(ms, 4), "cell_rhs_per_s": edge ** 3 / (ms * 1e-3)}


def amr_two_level_bench(steps):
    """The two-level reactor: 128^3 coarse level in uneven boxes (edges 8..64),
    GRF solid at porosity 0.5, 5 species, the fine level (ratio 2,
    subcycled) made from the coarse one where vfrac jumps; coarse steps of the
    hierarchy (mf_hier_step) timed after one untimed step"""
    import torch
    from mesoflow_amd import abi, amr, cases, runtime
    c = cases.porous_amr()
    s0 = c.make_state()
    ctx = runtime.Context(c)
    L0 = runtime.Level(ctx)
    L0.set_state(s0)
    tg, gr = cases.AMR_TAGS, cases.AMR_GRID
    tags, _ = amr.level_tag(L0, 0.0, tg["refine_phi"], tg["refine_phigrad"], tg["comps"])
    fb = amr.tags_to_boxes(tags, gr["n_error_buf"], gr["blocking_factor"], gr["max_grid_size"], gr["grid_eff"], 2,
                           periodic=(0, 1, 1))
    H = runtime.Hierarchy(L0, 2)
    H.regrid(fb)
    H.autotune()
    import ctypes as C

    def tiles_of(h):
        out_t = []
        for q in range(4):
            a, b, cc = C.c_int(), C.c_int(), C.c_int()
            runtime.check(runtime.lib().mf_level_get_tile(h, q, C.byref(a), C.byref(b), C.byref(cc)), "mf_level_get_tile")
            out_t.append("%d,%d,%d" % (a.value, b.value, cc.value))
        return ";".join(out_t)
    h1 = C.c_void_p()
    runtime.check(runtime.lib().mf_hier_level(H.h, 1, C.byref(h1)), "mf_hier_level")
    tiles = {"coarse": tiles_of(L0.h), "fine": tiles_of(h1)}
    nfine = int(sum(np.prod([h[d] - l[d] + 1 for d in range(3)]) for l, h in fb))
    ncrse = int(np.prod(c.n_cell))
    out = {"workload": "two-level porous reactor: 128^3 coarse level in %d uneven boxes (edges 8..64), GRF IB porosity "
                       "0.5, 5 species, 2 reactions, WENO5; fine level ratio 2 subcycled, %d boxes, %d cells "
                       "(%.1f%% of 256^3), from the vfrac tags" % (len(c.boxes), len(fb), nfine, 100.0 * nfine / 256 ** 3),
           "coarse_boxes": len(c.boxes), "fine_boxes": len(fb), "fine_cells": nfine, "steps": steps,
           "tiles": tiles}
    for name, m in (("batched", abi.MODE_BATCHED), ("graph", abi.MODE_GRAPH)):
        H.step(1, m)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        H.step(steps, m)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        out[name] = {"ms_per_coarse_step": el * 1e3, "cell_updates_per_s": (ncrse + 2 * nfine) / el}
        log("two-level %s: %.2f ms per coarse step" % (name, el * 1e3))
    H.close()
    L0.close()
    ctx.close()
    torch.cuda.empty_cache()
    return out


def amr_ops_bench(ctx, lvl, reps):
    """The two-level coupling operators and the tagging on the full level: the
    level's state as the fine data of a ratio-2 hierarchy (coarse = its
    average_down). Wall time of reps launches after one untimed launch, with
    algorithmic HBM bytes: average_down reads the fine and writes the coarse
    state; cell_cons_interp reads the coarse and writes the fine state;
    time interpolation reads two states and writes one; tagging reads the
    tagged components and writes one byte per cell."""
    import torch
    from mesoflow_amd import abi, amr, runtime
    nx, ny, nz = ctx.case.n_cell
    nv = ctx.nv
    fine = torch.empty((nv, nz, ny, nx), dtype=torch.float64, device="cuda")
    lvl.get_state(out=fine)
    cn = (nx // 2, ny // 2, nz // 2)
    crse = torch.zeros((nv, cn[2] + 2, cn[1] + 2, cn[0] + 2), dtype=torch.float64, device="cuda")
    other = fine.clone()
    dst = torch.empty_like(fine)
    tags = torch.zeros((nz, ny, nx), dtype=torch.uint8, device="cuda")
    pv = abi.Array4()
    runtime.check(runtime.lib().mf_level_state_view(lvl.h, pv), "mf_level_state_view")
    lvl.fill_ghosts(0.0)
    torch.cuda.synchronize()
    ctx.sync()
    fab_bytes = nv * nx * ny * nz * 8
    cfab_bytes = nv * cn[0] * cn[1] * cn[2] * 8
    full = ((0, 0, 0), (nx - 1, ny - 1, nz - 1))
    cbox = ((0, 0, 0), (cn[0] - 1, cn[1] - 1, cn[2] - 1))
    tagc = [abi.TEMP, abi.FLO_NVARS]

    def tag_call():
        ph, gr, cp, n = amr._tag_params([1e30, 1e30], [50.0, 1e-3], tagc)
        runtime.check(runtime.lib().mf_box_tag(ctx.h, abi.make_box(*full), pv, tags.data_ptr(), ph, gr, cp, n, 1),
                      "mf_box_tag")

    ops = {
        "average_down": (lambda: amr.average_down(ctx, cbox, (fine, (0, 0, 0)), (crse, (-1, -1, -1)), 2),
                         fab_bytes + cfab_bytes),
        "interp_cons": (lambda: amr.interp_cons(ctx, full, (crse, (-1, -1, -1)), (dst, (0, 0, 0)), 2),
                        fab_bytes + cfab_bytes),
        "time_interp": (lambda: amr.time_interp(ctx, full, (dst, (0, 0, 0)), (fine, (0, 0, 0)), (other, (0, 0, 0)),
                                                0.0, 1.0, 0.37), 3 * fab_bytes),
        "tag": (tag_call, len(tagc) * nx * ny * nz * 8 + nx * ny * nz),
    }
    out = {"level": "%dx%dx%d, nv %d, ratio 2" % (nx, ny, nz, nv), "reps": reps}
    for name, (fn, nbytes) in ops.items():
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        ms = (time.perf_counter() - t0) / reps * 1e3
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "algorithmic_GB": round(nbytes / 1e9, 3), "GB_per_s": round(gbs, 1),
                     "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["tagged_cells"] = int((tags != 0).sum().item())
    del fine, crse, other, dst, tags
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="porous256")
    ap.add_argument("--mode", default="batched", choices=["per_box", "batched", "graph"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile", default=None, help="tile shapes of the 4 kernel classes 'x,y,z;x,y,z;x,y,z;x,y,z' (skips the autotuner)")
    ap.add_argument("--calibrate", action="store_true", help="plain copy kernel for PMC scale factors")
    ap.add_argument("--compare-steps", type=int, default=3,
                    help="steps timed in each launch mode after the main region (0 = skip)")
    ap.add_argument("--split-substeps", type=int, default=2000,
                    help="chemistry substeps of the split-mode measurement (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the discrete-adjoint measurement at 128^3 (0 = skip)")
    ap.add_argument("--amr-reps", type=int, default=5,
                    help="launches of each two-level coupling operator and of the tagging at full size (0 = skip)")
    ap.add_argument("--amr-steps", type=int, default=5,
                    help="coarse steps of the two-level reactor (128^3 coarse, subcycled fine level; 0 = skip)")
    ap.add_argument("--box-rhs-reps", type=int, default=20,
                    help="calls of the per-box mf_box_rhs on a 32^3 box (0 = skip)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        raise SystemExit(2)

    import torch
    from mesoflow_amd import abi, runtime

    if args.calibrate:
        case, _ = make_case("sod64")
        ctx = runtime.Context(case)
        n = 1 << 28
        a = torch.zeros(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            runtime.check(runtime.lib().mf_util_copy(ctx.h, b.data_ptr(), a.data_ptr(), n), "mf_util_copy")
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "mf_util_copy", "bytes_read": n * 8, "bytes_written": n * 8}))
        return

    case, desc = make_case(args.config)
    nx, ny, nz = case.n_cell
    cells = nx * ny * nz
    nv = nv_of(case)
    log("config %s: %s" % (args.config, desc))
    t0 = time.time()
    state = case.make_state()
    log("initial state built in %.1f s (%.2f GB)" % (time.time() - t0, state.nbytes / 1e9))
    dev_state = torch.from_numpy(state).to("cuda")
    del state
    ctx = runtime.Context(case)
    lvl = runtime.Level(ctx)
    lvl.set_state(dev_state)
    del dev_state
    torch.cuda.synchronize()
    if args.tile:
        lvl.set_tiles([[int(v) for v in part.split(",")] for part in args.tile.split(";")])
    elif not args.no_autotune:
        tile = lvl.autotune()
        log("autotuned tile", tile)
    mode = {"per_box": abi.MODE_PER_BOX, "batched": abi.MODE_BATCHED, "graph": abi.MODE_GRAPH}[args.mode]
    if args.warmup:
        lvl.step(args.warmup, mode)
    torch.cuda.synchronize()

    # timed region: K steps, synchronised on both sides; kernels timed with HIP
    # events on the context stream (modes per_box / batched)
    t0 = time.perf_counter()
    if mode == abi.MODE_GRAPH:
        lvl.step(args.steps, mode)
        classes = None
    else:
        classes = lvl.step_timed(args.steps, mode)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = cells * args.steps / secs
    info = lvl.info()
    log("time %.6e dt %.3e step %d" % (info[0], info[1], info[2]))

    roof = None
    extra = {}
    step_bytes = 64.0 * nv * cells  # minimal fused RK3 stream per cell-update
    if classes:
        log("kernel classes (ms, launches):", json.dumps(classes))
        ms_faces, n_faces = classes["face_flux"]
        # dominant kernel: the face-flux sweep. One launch covers one direction of
        # one stage, i.e. 1/9 of a cell-update of the whole level.
        avg_s = ms_faces / max(n_faces, 1) * 1e-3
        alg_bytes = step_bytes / 9.0
        achieved = alg_bytes / avg_s / 1e9
        traffic, traffic_tile = read_traffic(args.config)
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_faces (WENO5 face-flux sweep)",
                "avg_launch_ms": avg_s * 1e3, "algorithmic_bytes_per_launch": alg_bytes,
                "traffic_source": ("rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE of profiles/traffic_%s.txt (tiles %s), "
                                   "scaled by the plain-copy calibration; not measured in this run"
                                   % (args.config, traffic_tile)) if traffic else None}
        tot = sum(v[0] for v in classes.values())
        extra["kernel_time_share"] = {k: round(v[0] / tot, 4) for k, v in classes.items()}
    extra["step_hbm_frac"] = (step_bytes * args.steps / secs / 1e9) / HBM_PEAK_GBS
    extra["tile"] = ";".join(",".join(str(v) for v in t) for t in lvl.tiles())
    nboxes = lvl.nboxes

    if args.compare_steps > 0:
        # the same steps launched per box (the reference's shape), as persistent
        # launches over the tile list, and replayed from a hipGraph; no events
        modes = {}
        for name, m in (("per_box", abi.MODE_PER_BOX), ("batched", abi.MODE_BATCHED), ("graph", abi.MODE_GRAPH)):
            lvl.step(1, m)
            torch.cuda.synchronize()
            ta = time.perf_counter()
            lvl.step(args.compare_steps, m)
            torch.cuda.synchronize()
            modes[name] = (time.perf_counter() - ta) / args.compare_steps * 1e3
            log("mode %s: %.3f ms/step" % (name, modes[name]))
        extra["ms_per_step_by_mode"] = modes
        extra["batched_speedup_vs_per_box"] = modes["per_box"] / modes["batched"]
        extra["graph_speedup_vs_per_box"] = modes["per_box"] / modes["graph"]

    if args.split_substeps > 0:
        # split-chemistry mode (Evolve_split): one flow-only step gives the
        # chemistry phase its lag state, then one phase of split_substeps
        # explicit-midpoint substeps of every cell, fused in one kernel per
        # 1024 substeps
        t_now, dt_now = lvl.info()[:2]
        lvl.evolve(t_now + 0.5 * dt_now, only_flow=True)
        torch.cuda.synchronize()
        dt_rk = 1.0e-6
        nsub = substeps(args.split_substeps * dt_rk, dt_rk)
        ta = time.perf_counter()
        lvl.advance_chemistry(0.0, args.split_substeps * dt_rk, dt_rk)
        torch.cuda.synchronize()
        secs_c = time.perf_counter() - ta
        sc = {"substeps": nsub, "dt_rk": dt_rk, "ms": secs_c * 1e3, "cell_substeps_per_s": cells * nsub / secs_c}
        if not args.no_cpu_baseline:
            sc["cpu_cell_substeps_per_s"] = split_chem_cpu(args.config)
            sc["cpu_sample"] = "oracle chemistry phase on the 64^3 sample, 100 substeps, OMP threads %s" % os.environ.get(
                "OMP_NUM_THREADS", os.cpu_count())
        extra["split_chemistry"] = sc
        log("split chemistry: %s" % json.dumps(sc))

    if args.box_rhs_reps > 0:
        extra["box_rhs"] = box_rhs_bench(ctx, lvl, args.box_rhs_reps)
        log("mf_box_rhs: %s" % json.dumps(extra["box_rhs"]))

    if args.amr_reps > 0:
        extra["amr_ops"] = amr_ops_bench(ctx, lvl, args.amr_reps)
        log("amr operators: %s" % json.dumps(extra["amr_ops"]))

    if args.adjoint_steps > 0 or args.amr_steps > 0:
        lvl.close()
        ctx.close()
        lvl = ctx = None
        torch.cuda.empty_cache()
    if args.amr_steps > 0:
        log("two-level reactor ...")
        extra["two_level"] = amr_two_level_bench(args.amr_steps)
    if args.adjoint_steps > 0:
        log("discrete adjoint at 128^3 ...")
        extra["adjoint"] = adjoint_bench(args.adjoint_steps)

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a sample ...")
        cpu = cpu_baseline(args.config, args.cpu_budget)
        log("cpu baseline", cpu)

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded initial state of the named workload (no datasets)",
        "config": {"workload": desc, "cells": cells, "nspec": case.nspec, "nv": nv, "order_hyp": int(case.params.order_hyp),
                   "mode": args.mode, "boxes": nboxes},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(out), flush=True)
    if lvl is not None:
        lvl.close()
        ctx.close()


if __name__ == "__main__":
    main()
