"""Throughput of the MI355X etch-loop hot path on the single-flux-step-at-scale workload.

Workload (one GPU): pre-carved trench 2 um wide and 20 um deep (aspect ratio 10) with a 2 um
mask, dx = 0.1 um, level set 160 x 40 x 260 float32; per etch step: marching cubes -> LBVH ->
coupled flux with 8M power-cosine ions (exponent 1000, 100 eV) and 5 flux/coverage fixed-point
iterations of 8M F + 8M O Russian-roulette neutrals (88M primary rays per step) -> Belen rate ->
velocity extension -> CFL substeps of upwind advection -> FSM reinit. A step is one such pass;
the surface keeps etching from step to step.

    python bench.py [--steps K] [--warmup W] [--no-cpu-baseline] [--no-small-hole] [--no-batch]

Prints one JSON line: steps/s (value), rays/s, ray-hits/s, the roofline of the dominant kernel
(the neutral trace) measured with HIP events on the launch stream, the CPU baseline (the
oracle's C/OpenMP restatement timed on a bounded sample of the same workload), and beside them
the wall time of the reference's own small-hole loop (scripts/vps_sweep.py) against its
published RTX 3090 figure, and the batched configuration (64 features in lockstep, one launch
per trace) against the same features run one after the other.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "plasma-etching-code_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "3D etch steps/s & rays/s (1 feature, 64-batch); fwd+adjoint / fwd cost ratio"
HBM_PEAK_GBS = 8000.0
# Each k_trace_neutral_rr launch carries the F and the O rays of one fixed-point iteration
# (2 x 8M primary rays); a ray-hit of either species reads its 36 B triangle and 4 B bare value
# and read-modify-writes its 4 B accumulator.
NEUTRAL_BYTES_PER_HIT = 48.0     # 36 B triangle + 4 B bare + 8 B accumulator read-modify-write
ION_BYTES_PER_HIT = 76.0         # 36 B triangle + 5 x 8 B accumulators

# workload: the AR-10 trench of the single-flux-step-at-scale configuration
W = dict(Lx=16.0, Ly=4.0, Lz=26.0, dx=0.1, width=2.0, floor=2.0, sub_top=22.0, mask_th=2.0)
N_ION = 8_000_000
N_NEU = 8_000_000           # per species per fixed-point iteration (F and O: 16M per iteration)
N_FP = 5
DT = 0.075


def workload_par():
    from petch import PAR
    from petch.api import SF6O2
    m = SF6O2()
    par = dict(m.par)
    par.update(periodic_y=1, n_fp=N_FP, ion_source='powcos', ion_exponent=1000.0, Emean=100.0)
    return par, m.flags


def build_loop(n_steps, timer=None, autotune=True):
    import torch
    from petch import Flags, ops
    from petch.threed import EtchLoop, grid_shape
    dev = torch.device("cuda", 0)
    nx, ny, nz = grid_shape(W['Lx'], W['Ly'], W['Lz'], W['dx'])
    # carved trench: solid below the floor and beside the 2 um slot up to the mask top
    phi, _ = ops.init_sdf(nx, ny, nz, W['dx'], W['Lx'], W['Ly'], W['width'],
                          W['sub_top'] + W['mask_th'] - W['floor'], W['floor'], False, device=dev)
    _, mask = ops.init_sdf(nx, ny, nz, W['dx'], W['Lx'], W['Ly'], W['width'], W['mask_th'],
                           W['sub_top'], False, device=dev)
    import numpy as np
    geo = dict(xs=np.arange(nx) * W['dx'], ys=np.arange(ny) * W['dx'], zs=np.arange(nz) * W['dx'],
               dx=W['dx'], Lx=W['Lx'], Ly=W['Ly'], Lz=W['Lz'], sub_top=W['sub_top'],
               trench_width=W['width'], hole=False)
    par, fl = workload_par()
    flags = Flags(**fl)
    # the first (untimed) step picks the BVH build and the trace occupancy by measurement
    loop = EtchLoop(phi, mask, geo, W['mask_th'], DT * n_steps, n_steps, par, flags, N_ION, N_NEU,
                    reinit_method="fsm", timer=timer, autotune=autotune)
    return loop, (nx, ny, nz)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_neutral_rr_trench_ar10_traffic.txt")


def measured_traffic():
    """HBM bytes per k_trace_neutral_rr launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/ (KiB per dispatch, each scaled by what the same counter reads for
    a plain device copy of known size). None when no summary is present."""
    try:
        vals = {}
        for ln in open(TRAFFIC_FILE):
            if ":" in ln and not ln.startswith("#"):
                k, v = ln.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
        fetch = vals["fetch_kib_per_dispatch"] * vals["fetch_scale"]
        write = vals["write_kib_per_dispatch"] * vals["write_scale"]
        return 1024.0 * (fetch + write), os.path.relpath(TRAFFIC_FILE, ROOT)
    except (OSError, KeyError):
        return None, None


def _cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs visible)"


def cpu_baseline(scale=1):
    """Oracle (C/OpenMP traces + FP64 numpy level set) on one step of the same workload; with
    scale > 1 the ray counts are divided by `scale` and the ray-proportional time scaled back."""
    import numpy as np
    from oracle import flux as ofl
    from oracle import levelset as ol
    from oracle import chem as oc
    from oracle import native as on
    from petch import Flags
    par, fl = workload_par()
    flags = Flags(**fl)
    L = W
    phi = ol.init_sdf(L['Lx'], L['Ly'], L['Lz'], L['dx'], L['width'],
                      L['sub_top'] + L['mask_th'] - L['floor'], L['floor'], False).astype(np.float32)
    mask = ol.trench_mask(L['Lx'], L['Ly'], L['Lz'], L['dx'], L['width'], L['mask_th'],
                          L['sub_top'], False)
    geo = dict(Lx=L['Lx'], Ly=L['Ly'], Lz=L['Lz'], dx=L['dx'], sub_top=L['sub_top'])
    t0 = time.time()
    v, f = on.marching_cubes(phi, L['dx'])
    M = ofl.OracleMesh(v, f)
    t_mesh = time.time() - t0
    t0 = time.time()
    out = ofl.coupled_flux(M, geo, par, N_ION // scale, N_NEU // scale, 0, flags)
    t_flux = time.time() - t0
    t0 = time.time()
    is_mask = oc.faces_in_mask(M.cen, geo, L['mask_th'], L['width'], hole=False)
    V = np.nan_to_num(oc.surface_rate(out['m_i'], out['m_F'], out['m_O'], out['cos_i'], is_mask,
                                      par, flags, ion_yield=out['ion_yield']))
    Fs = ofl.extend(M, V.astype(np.float32), phi, L['dx'], 4 * L['dx'])
    vmx = max(float(V.max()), 1e-6)
    nsub = max(1, min(int(np.ceil(vmx * DT / (0.4 * L['dx']))), 160))
    p = ol.advect_pinned(phi.astype(np.float64), Fs, mask, phi.astype(np.float64), L['dx'],
                         DT / nsub, nsub)
    ol.reinit_fsm(p, L['dx'], 6 * L['dx'])
    t_rest = time.time() - t0
    step_s = t_mesh + t_rest + t_flux * scale
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=1.0 / step_s, unit="steps/s", cores=cores, kind="port", cpu=_cpu_model(),
                sample=(f"one etch step of the same trench state on the oracle: marching cubes, "
                        f"rate, extension, {nsub} advection substeps and FSM reinit at full size "
                        f"({t_mesh + t_rest:.1f} s) plus the coupled flux with "
                        f"{'all' if scale == 1 else f'1/{scale} of'} the rays ({t_flux:.1f} s"
                        f"{'' if scale == 1 else f', scaled x{scale}'}); traces in C with OpenMP, "
                        f"level set in single-threaded numpy"))


# The reference's own timed loop (scripts/vps_sweep.py:46-55): hole of diameter 6 um, extent 14,
# dx 0.25 (56 x 56 x 96), 40 steps over 3.0 s, 30k ion + 30k neutral rays per species, n_fp 1,
# betaE 0.7, coverage sticking + warm start, fsm reinit; published wall time 1.61 s on an
# RTX 3090 (box_results/vps_sweep_result.json:16-20), i.e. 24.8 steps/s.
SMALL_HOLE_REF_WALL_S = 1.6116929054260254


def small_hole_run(n_steps=40, rate=0.15):
    import petch
    from petch import threed as t3
    p = dict(petch.PAR)
    p['rate_scale'] = rate
    p['betaE'] = 0.7
    p.update(flux_smooth_gpu=True, gpu_source=True, gpu_mesh=True, gpu_warmstart=True,
             device_flux=True, n_fp=1)
    geo = dict(Lx=14.0, Ly=14.0, Lz=24, mask_th=2, sub_top=18, hole=True, t_end=3.0)
    flags = petch.Flags(coverage_sticking=True, sampling="sobol", warm_start_coverage=True)
    t0 = time.time()
    g = t3.run_etch_3d(trench_width=6.0, dx=0.25, n_steps=n_steps, par=p, flags=flags,
                       n_ion=30000, n_neu=30000, reinit_method="fsm", verbose=False, **geo)
    return time.time() - t0, t3.center_depth_3d(g)


def small_hole():
    """Whole-call wall time of the reference's own small-hole loop, for comparison with its
    published figure (not the headline value: that loop is launch-bound, not ray-bound)."""
    small_hole_run(n_steps=4)                      # warm-up, as the reference's script does
    walls, depth = [], None
    for _ in range(3):
        w, depth = small_hole_run()
        walls.append(w)
    w = sorted(walls)[1]
    return {"workload": "hole d 6 um, 56x56x96, 40 steps, 30k rays/species (scripts/vps_sweep.py)",
            "wall_s": w, "steps_per_s": 40.0 / w, "center_depth_um": depth,
            "reference_wall_s_rtx3090": SMALL_HOLE_REF_WALL_S,
            "speedup_vs_reference": SMALL_HOLE_REF_WALL_S / w}


def batched(n_feat=64, n_steps=40, sequential=True):
    """The batched configuration, whole runs (40 steps over 3.0 s): all features in lockstep with
    one launch per trace (petch.batch) against the same features run one after the other through
    run_etch_3d. Same streams, so the depths must agree. sequential=False times the batched run
    only (for profiling it)."""
    import torch
    from petch import batch as pb
    from petch import threed as t3
    feats = pb.seeded_features(n_feat)
    t_end = 3.0 * n_steps / 40.0
    pb.run_etch_3d_batch(pb.seeded_features(4, seed=99), t_end=t_end, n_steps=2)  # warm-up
    torch.cuda.synchronize()
    t0 = time.time()
    gb = pb.run_etch_3d_batch(feats, t_end=t_end, n_steps=n_steps)
    torch.cuda.synchronize()
    wb = time.time() - t0
    if not sequential:
        lb = gb[0]['timings']['total']
        return {"workload": f"{n_feat} seeded holes/trenches, {n_steps} steps over {t_end:g} s",
                "batched_steps_s": lb, "features_per_s_batched": n_feat / lb,
                "batched_wall_s": wb}
    t0 = time.time()
    gs = [t3.run_etch_3d(**ft, t_end=t_end, n_steps=n_steps, reinit_method="fsm", verbose=False)
          for ft in feats]
    torch.cuda.synchronize()
    ws = time.time() - t0
    dd = max(abs(t3.max_depth_3d(a) - t3.max_depth_3d(b)) for a, b in zip(gb, gs))
    # the etch steps alone (run_etch_3d's timings['total']: the step loop, without building the
    # initial level sets and copying phi back, which both paths do once per feature)
    lb = gb[0]['timings']['total']
    ls = sum(g['timings']['total'] for g in gs)
    return {"workload": f"{n_feat} seeded holes/trenches (d U(2,10) um, dx 0.25, 48^2..160^2 x "
                        f"80..200 cells, 400k rays/feature/step), {n_steps} steps over "
                        f"{t_end:g} s",
            "batched_steps_s": lb, "sequential_steps_s": ls,
            "features_per_s_batched": n_feat / lb,
            "features_per_s_sequential": n_feat / ls,
            "speedup": ls / lb, "batched_wall_s": wb, "sequential_wall_s": ws,
            "wall_speedup": ws / wb, "max_depth_diff_um": dd}


def adjoint_cost(nsub=3, reps=10):
    """Cost of the differentiable advection (petch.adjoint.advect) on the trench grid: the loss
    (w . phi after nsub substeps) against the loss and its gradient (the HIP adjoint through
    every substep)."""
    import torch
    from petch import adjoint, ops
    from petch.threed import grid_shape
    dev = torch.device("cuda", 0)
    nx, ny, nz = grid_shape(W['Lx'], W['Ly'], W['Lz'], W['dx'])
    phi, mask = ops.init_sdf(nx, ny, nz, W['dx'], W['Lx'], W['Ly'], W['width'],
                             W['sub_top'] + W['mask_th'] - W['floor'], W['floor'], False,
                             device=dev)
    g = torch.Generator(device="cpu").manual_seed(3)
    F = torch.where(phi.abs() < 4 * W['dx'], torch.rand(phi.shape, generator=g).to(dev), 0.0)
    mask_phi = phi.clone()
    w = torch.rand(phi.shape, generator=g).to(dev)
    dt = 0.4 * W['dx'] / nsub

    def run(grad):
        x = phi.clone().requires_grad_(grad)
        f = F.clone().requires_grad_(grad)
        loss = (adjoint.advect(x, f, mask, mask_phi, W['dx'], dt, nsub) * w).sum()
        if grad:
            loss.backward()

    res = {}
    for grad in (False, True):
        run(grad)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run(grad)
        e.record()
        torch.cuda.synchronize()
        res["fwd_adj_ms" if grad else "fwd_ms"] = s.elapsed_time(e) / reps
    return {"scope": f"level-set advection, {nsub} substeps on {nx}x{ny}x{nz}; fwd = the loss "
                     "w . phi, fwd_adj = the loss and its gradient in phi and F",
            **res, "ratio": res["fwd_adj_ms"] / res["fwd_ms"]}


def radiosity_adjoint_cost(reps=5, n_ff=64, tol=1e-10):
    """Cost of the differentiable radiosity solve (petch.adjoint.radiosity) on the trench
    surface: the converged solve (tol) forward against forward + transposed solve, with the
    form factors of n_ff rays per face assembled once outside the timing."""
    import torch
    from petch import adjoint, ops
    from petch.flux import SurfaceMesh
    from petch.threed import grid_shape
    dev = torch.device("cuda", 0)
    nx, ny, nz = grid_shape(W['Lx'], W['Ly'], W['Lz'], W['dx'])
    phi, _ = ops.init_sdf(nx, ny, nz, W['dx'], W['Lx'], W['Ly'], W['width'],
                          W['sub_top'] + W['mask_th'] - W['floor'], W['floor'], False, device=dev)
    mesh = SurfaceMesh(dev).build(phi, W['dx'])
    gn = ops.gas_normals(phi, W['dx'], mesh.centroids, mesh.normals)
    box = ops.make_box(W['Lx'], W['Ly'], W['Lz'], W['Lz'] - W['dx'], 1)
    hit = ops.trace_ff(mesh.bvh, box, mesh.centroids, gn, n_ff, 5)
    nf = mesh.nf
    row_ptr, col, D = ops.radiosity_assemble(hit, nf, n_ff)
    g = torch.Generator(device="cpu").manual_seed(4)
    bare = (0.2 + 0.8 * torch.rand(nf, generator=g)).to(dev)
    w = torch.rand(nf, generator=g).to(dev)

    def run(grad):
        x = bare.clone().requires_grad_(grad)
        loss = (adjoint.radiosity(x, D[:nf], hit, row_ptr, col, n_ff, 0.7, 0, tol) * w).sum()
        if grad:
            loss.backward()

    res = {}
    for grad in (False, True):
        run(grad)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run(grad)
        e.record()
        torch.cuda.synchronize()
        res["fwd_adj_ms" if grad else "fwd_ms"] = s.elapsed_time(e) / reps
    return {"scope": f"radiosity solve to tol {tol:g} ({nf} faces, {n_ff} form-factor rays each)",
            **res, "ratio": res["fwd_adj_ms"] / res["fwd_ms"]}


def neutral_adjoint_cost(reps=3, beta=0.7):
    """Cost of the differentiable RR neutral trace (petch.adjoint.neutral_rr) on the trench
    surface with the bench's 8M rays of one species: the trace forward against forward +
    path-replay adjoint (every roulette segment of two or more hits traced twice more)."""
    import torch
    from petch import adjoint, ops
    from petch.flux import SurfaceMesh
    from petch.threed import grid_shape
    dev = torch.device("cuda", 0)
    nx, ny, nz = grid_shape(W['Lx'], W['Ly'], W['Lz'], W['dx'])
    phi, _ = ops.init_sdf(nx, ny, nz, W['dx'], W['Lx'], W['Ly'], W['width'],
                          W['sub_top'] + W['mask_th'] - W['floor'], W['floor'], False, device=dev)
    mesh = SurfaceMesh(dev, rays_per_step=N_NEU * 2 * N_FP).build(phi, W['dx'])
    box = ops.make_box(W['Lx'], W['Ly'], W['Lz'], W['Lz'] - W['dx'], 1)
    src = ops.make_source(ops.SRC_NEUTRAL)
    g = torch.Generator(device="cpu").manual_seed(6)
    bare = (0.2 + 0.8 * torch.rand(mesh.nf, generator=g)).to(dev)
    w = torch.randn(mesh.nf, generator=g).to(dev)

    def run(grad):
        x = bare.clone().requires_grad_(grad)
        loss = (adjoint.neutral_rr(mesh.bvh, box, src, x, beta, N_NEU, 11) * w).sum()
        if grad:
            loss.backward()

    res = {}
    for grad in (False, True):
        run(grad)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run(grad)
        e.record()
        torch.cuda.synchronize()
        res["fwd_adj_ms" if grad else "fwd_ms"] = s.elapsed_time(e) / reps
    return {"scope": f"RR neutral trace, {N_NEU} rays of one species on the trench surface "
                     f"({mesh.nf} faces), path-replay adjoint to bare",
            **res, "ratio": res["fwd_adj_ms"] / res["fwd_ms"]}


def step_adjoint_cost(reps=2, nsub=3):
    """Cost of one differentiable etch step on the trench surface of the workload, from the mean
    ion energy to a loss on the advanced level set: the faithful ion trace (8M ions), N_FP rounds
    of F and O RR traces (8M rays each) with finish and Belen updates, the rate, the extension
    and nsub advection substeps. fwd = the loss; fwd_adj = the loss and d loss / dE (forward-mode
    ion trace, path-replay RR adjoints, the closed-form and transposed stages). The geometry (MC,
    BVH) is built once outside the timing and the reinit is not part of it."""
    import torch
    from petch import adjoint, ops
    from petch.chemistry import chem_params
    from petch.flux import SurfaceMesh
    from petch.params import Flags
    from petch.threed import grid_shape
    dev = torch.device("cuda", 0)
    nx, ny, nz = grid_shape(W['Lx'], W['Ly'], W['Lz'], W['dx'])
    phi, _ = ops.init_sdf(nx, ny, nz, W['dx'], W['Lx'], W['Ly'], W['width'],
                          W['sub_top'] + W['mask_th'] - W['floor'], W['floor'], False, device=dev)
    _, mask = ops.init_sdf(nx, ny, nz, W['dx'], W['Lx'], W['Ly'], W['width'], W['mask_th'],
                           W['sub_top'], False, device=dev)
    mesh = SurfaceMesh(dev, rays_per_step=N_NEU * 2 * N_FP).build(phi, W['dx'])
    geo = dict(Lx=W['Lx'], Ly=W['Ly'], Lz=W['Lz'], dx=W['dx'])
    par, fl = workload_par()
    flags = Flags(**fl)
    chem = chem_params(par, flags)
    mg = ops.MaskGeom(float(W['sub_top']), float(W['mask_th']), float(W['Lx']), float(W['Ly']),
                      float(W['width']), 0)
    iy_rate = getattr(flags, "chemistry", "") == "belen"
    w = torch.randn(phi.shape, generator=torch.Generator(device="cpu").manual_seed(8)).to(dev)
    band = 4 * W['dx']                       # the etch loop's extension band

    def run(grad):
        E = torch.tensor(float(par['Emean']), dtype=torch.float64, requires_grad=grad)
        m_i, cos_i, iy = adjoint.ion_terms(mesh, geo, par, N_ION, 1, E)
        _, m_F, m_O = adjoint.coupled_coverage(mesh, geo, par, flags, m_i, cos_i, iy, N_NEU, 1,
                                               n_fp=N_FP)
        phi1 = adjoint.advance(mesh, phi, mask, phi, W['dx'], DT, nsub, band, m_i, m_F, m_O,
                               cos_i, iy if iy_rate else None, mg, chem)
        loss = (phi1 * w).sum()
        if grad:
            loss.backward()
            return float(E.grad)
        return None

    res = {}
    for grad in (False, True):
        run(grad)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            g = run(grad)
        e.record()
        torch.cuda.synchronize()
        res["fwd_adj_ms" if grad else "fwd_ms"] = s.elapsed_time(e) / reps
        if grad:
            res["dloss_dE"] = g
    return {"scope": f"one etch step from the ion energy on the trench surface ({mesh.nf} faces): "
                     f"{N_ION} ions, {N_FP} x ({N_NEU} F + {N_NEU} O) RR neutrals, rate, "
                     f"extension, {nsub} advection substeps; geometry fixed, no reinit",
            **res, "ratio": res["fwd_adj_ms"] / res["fwd_ms"]}


def gradient_cost(reps=2, n_rays=200_000, checkpoint_every=5):
    """Cost of the differentiable 20-step hole (petch.diffetch, deterministic mode): the small
    hole at dx 0.2 um, 3.0 s over 20 steps, 200k ions and 200k F + 200k O neutrals per round;
    fwd = the profile loss of the final level set, fwd_adj = the loss and its gradient in
    (Emean, Fflux, Oflux, rate_scale), ckpt = the same with segments of 5 steps recomputed in the
    backward pass. Each is run once untimed first."""
    import torch
    from petch import diffetch as de, ops
    from petch.api import SF6O2
    from petch.params import Flags
    ops.set_deterministic(True)
    try:
        m = SF6O2()
        case = dict(Lx=14.0, Ly=14.0, Lz=24.0, dx=0.2, trench_width=6.0, mask_th=2.0,
                    sub_top=18.0, hole=True, t_end=3.0, n_steps=20)
        dd = de.DiffEtch(**case, par=dict(m.par), flags=Flags(**m.flags), n_ion=n_rays,
                         n_neu=n_rays)
        tg = de.DiffEtch(**case, par=dict(m.par), flags=Flags(**m.flags), n_ion=n_rays,
                         n_neu=n_rays, seed_offset=1234)
        target = de.profile_target(tg, tg.theta0 * torch.tensor([1.1, 0.9, 1.1, 1.1],
                                                                dtype=torch.float64))
        zsel = de.zsel_feature(dd, below=8.0)
        runs = {"fwd_s": lambda: de.loss_only(dd, dd.theta0, target, zsel),
                "fwd_adj_s": lambda: de.loss_and_gradient(dd, dd.theta0, target, zsel),
                "ckpt_s": lambda: de.loss_and_gradient(dd, dd.theta0, target, zsel,
                                                       checkpoint_every=checkpoint_every)}
        res = {}
        for key, fn in runs.items():
            fn()
            torch.cuda.synchronize()
            t = time.time()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            res[key] = (time.time() - t) / reps
        _, g = de.loss_and_gradient(dd, dd.theta0, target, zsel)
    finally:
        ops.set_deterministic(False)
    return {"scope": "20-step differentiable hole (dx 0.2 um, 70x70x120, 200k ions + 2 x 200k "
                     "neutrals per round), profile loss, gradient in Emean, Fflux, Oflux, "
                     "rate_scale; deterministic mode",
            **res, "ratio": res["fwd_adj_s"] / res["fwd_s"],
            "ratio_checkpointed": res["ckpt_s"] / res["fwd_s"],
            "gradient": [float(x) for x in g]}


def config4_cost(n_variants=8, n_rays=1_000_000):
    """The batched differentiable configuration (petch.diffetch.CONFIG4: a 60 x 10 x 240 trench,
    30 steps over 6 s, 1M rays per species and step, five-parameter gradient of each variant's
    profile loss, checkpoints every 6 steps), run variant after variant in deterministic mode;
    seconds per variant for its target run plus its checkpointed forward and adjoint, and the
    same variant's plain forward for the cost ratio."""
    import torch
    from petch import diffetch as de, ops
    from petch.params import Flags
    ops.set_deterministic(True)
    try:
        de.variant_gradients(1, n_rays=n_rays,
                             case=dict(de.CONFIG4, n_steps=2, t_end=de.CONFIG4['t_end'] / 15))
        def progress(done, secs):
            if done % 16 == 0:
                print(f"bench: config4 {done}/{n_variants} variants, {secs:.1f} s",
                      file=sys.stderr, flush=True)
        res, secs = de.variant_gradients(n_variants, n_rays=n_rays, progress=progress)
        (par, flags), = de.variant_recipes(1)
        dd = de.DiffEtch(**de.CONFIG4, par=par, flags=Flags(**flags), n_ion=n_rays, n_neu=n_rays,
                         params=de.PARAMS5)
        target = de.profile_target(dd, dd.theta0)
        zsel = de.zsel_feature(dd)
        torch.cuda.synchronize()
        t = time.time()
        de.loss_only(dd, dd.theta0, target, zsel)
        torch.cuda.synchronize()
        fwd = time.time() - t
        t = time.time()
        de.loss_and_gradient(dd, dd.theta0, target, zsel, checkpoint_every=6)
        torch.cuda.synchronize()
        ckpt = time.time() - t
    finally:
        ops.set_deterministic(False)
    return {"scope": f"{n_variants} of the 256 variants of the 60x10x240 trench, 30 steps, "
                     f"{n_rays} rays per species and step, 5-parameter gradient, checkpoints "
                     f"every 6 steps, variants one after the other, deterministic mode",
            "variants": n_variants, "seconds": secs, "seconds_per_variant": secs / n_variants,
            "fwd_s": fwd, "fwd_adj_checkpointed_s": ckpt, "ratio_checkpointed": ckpt / fwd,
            "losses": [float(l) for l, _ in res],
            "gradient_first": [float(x) for x in res[0][1]]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    # two warm-up steps: the first runs from full coverage, the second is the first with the
    # coverage warm start, whose first torch ops load their kernels (stalls of ~0.3 s)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-small-hole", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run plain device copies of 256 MiB (reference for PMC byte counts)")
    ap.add_argument("--batch-only", type=int, default=0, metavar="STEPS",
                    help="only run the batched 64-feature etch for STEPS steps (one JSON line; "
                         "for profiling)")
    ap.add_argument("--config4", type=int, default=0, metavar="N",
                    help="only run N variants of the batched differentiable trench (one JSON "
                         "line)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time the differentiable pieces (one JSON line; for profiling)")
    a = ap.parse_args(argv)
    if a.config4:
        print(json.dumps({"config4": config4_cost(a.config4)}), flush=True)
        return
    if a.batch_only:
        print(json.dumps({"batch": batched(n_steps=a.batch_only, sequential=False)}), flush=True)
        return
    if a.adjoint_only:
        print(json.dumps({"adjoint": adjoint_cost(), "adjoint_radiosity": radiosity_adjoint_cost(),
                          "adjoint_neutral": neutral_adjoint_cost(),
                          "adjoint_step": step_adjoint_cost(),
                          "gradient": gradient_cost()}), flush=True)
        return
    if a.copy_probe:
        import torch
        x = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        x.fill_(1.0)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": x.numel() * 4, "copy_probe_bytes_written": x.numel() * 4,
                          "dispatches": 5}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the etch step of one feature runs on one device "
              "(batches of features would be split across GPUs, one process each, with no "
              "communication inside a step); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from petch import ops
    from petch.profiling import StageTimer
    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    timer = StageTimer(enabled=True)
    total = a.warmup + a.steps
    loop, shape = build_loop(total, timer)
    print(f"bench: trench {shape} n_ion={N_ION} n_neu={N_NEU}x2x{N_FP}", file=sys.stderr, flush=True)
    for s in range(a.warmup):
        t = time.time()
        loop.step(s)
        torch.cuda.synchronize()
        print(f"bench: warmup step {s} {time.time() - t:.2f}s faces={loop.last['nf']}",
              file=sys.stderr, flush=True)
    timer.resolve()
    timer.reset()
    ops.trace_hit_counts(reset=True)
    torch.cuda.synchronize()
    t0 = time.time()
    faces = []
    for s in range(a.warmup, total):
        info = loop.step(s)
        faces.append(info['nf'])
        print(f"bench: step {s} faces={info['nf']} nsub={info['nsub']}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.time() - t0
    hits_n, hits_i = ops.trace_hit_counts(reset=True)
    timer.resolve()
    ms_step = 1000.0 * wall / a.steps
    rays_step = N_ION + N_FP * 2 * N_NEU
    nt_ms = timer.ms["neutral_trace"]
    nt_cnt = max(timer.count["neutral_trace"], 1)
    avg_launch_ms = nt_ms / nt_cnt
    bytes_launch = NEUTRAL_BYTES_PER_HIT * hits_n / nt_cnt
    achieved = bytes_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic()
    res = {
        "metric": METRIC,
        "value": a.steps / wall,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single flux step at scale: AR-10 trench (W 2 um, depth 20 um, "
                               "2 um mask, dx 0.1 um, 160x40x260), 8M power-cosine ions "
                               "(n=1000, 100 eV) + 5 fixed-point iterations x (8M F + 8M O) "
                               "Russian-roulette neutrals per step, periodic lateral BC",
                   "grid": list(shape), "faces_mean": sum(faces) / len(faces),
                   "rays_per_step": rays_step},
        "rays_per_s": rays_step * a.steps / wall,
        "ray_hits_per_s": (hits_n + hits_i) / wall,
        "stages_ms_per_step": {k: v / a.steps for k, v in sorted(timer.ms.items())},
        "roofline": {"kernel": "k_trace_neutral_rr", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "avg_launch_ms": avg_launch_ms,
                     "hits_per_launch": hits_n / nt_cnt,
                     "bytes_per_hit": NEUTRAL_BYTES_PER_HIT},
    }
    if loop.tuner is not None and loop.tuned is not None:
        res["autotune"] = {"bvh_quality": loop.tuned[0], "trace_blocks_per_cu": loop.tuned[1],
                           "candidates_ms": {f"q{q}/occ{o}": [round(b, 3), round(t, 3)]
                                             for (q, o), (b, t) in loop.tuner.last.items()},
                           "note": "build ms, paired neutral trace ms (occupancy 0 = the "
                                   "kernel's limit)"}
    ops.set_trace_occupancy(0)
    if not a.no_small_hole:
        res["small_hole"] = small_hole()
    if not a.no_batch:
        res["batch"] = batched()
        res["adjoint"] = adjoint_cost()
        res["adjoint_radiosity"] = radiosity_adjoint_cost()
        res["adjoint_neutral"] = neutral_adjoint_cost()
        res["adjoint_step"] = step_adjoint_cost()
        res["gradient"] = gradient_cost()
        res["config4"] = config4_cost()
    if not a.no_cpu_baseline:
        try:
            res["cpu_baseline"] = cpu_baseline()
        except Exception as e:                              # reported, not fatal
            res["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
