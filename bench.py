"""Benchmark: Stokes solves/s to relative residual 1e-6 on one MI355X.

One step = one complete Stokes solve of the 1,013,958-velocity-dof shell
(level-4 icosahedral surface x 16 layers, 245,760 P2 tets, 3-level MG):
element assembly with rotation and BC elimination, element-local Galerkin
coarse operators, coarse Cholesky factorisation, Patera pressure iteration
with the MGPCG velocity solves, until
    rel_res = max(|r_P| / |r_P^cold|, |r_U| / |f_U|) <= 1e-6
(RMS norms; r_P^cold is the pressure residual of the velocity solve at P=0,
the same denominator for every form).  Inputs are seeded synthetic fields
(temperature with spherical-harmonic perturbation, viscosity contrast 1e4,
6 Voronoi plates at the surface, free slip at the CMB).

Prints one JSON line.  Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1]
                  [--form fp64|mixed] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def sweep_bytes(nnzb, nnod, val_bytes=8, vec_bytes=8):
    """Algorithmic bytes of one fused Jacobi sweep r = r0 - K d; d' = d + w D^-1 r
    on the fine level: every stored 3x3 block once (values + int32 column),
    the row table, and per row the vectors d (gather, perfect reuse), r0,
    D^-1 read and d' written."""
    return nnzb * (9 * val_bytes + 4) + nnod * 4 + nnod * 3 * 4 * vec_bytes


def cpu_baseline(ctx, mesh, inp, n_cg, seed=0):
    """Oracle (NumPy/SciPy, one core) timed on a bounded sample of one solve."""
    from threadpoolctl import threadpool_limits
    import scipy.sparse.linalg as spla
    from oracle import assembly as OA, elements as OE
    from m3tet_sph_amd import api
    with threadpool_limits(limits=1):
        rng = np.random.default_rng(seed)
        nsamp = 2048
        T = OE.sf_tables(14)
        t0 = time.perf_counter()
        for kind in OA.CLASS_NAMES:
            els = mesh.classes[kind]
            k = max(1, int(round(nsamp * len(els) / mesh.nel)))
            if len(els):
                OA.class_kernel_opt(kind, rng.choice(els, size=min(k, len(els)), replace=False), mesh,
                                    inp["Visc"], inp["Dens"], inp["phys"], T)
        t_asm = (time.perf_counter() - t0) * mesh.nel / nsamp
        K0 = api.get_operator(ctx, 0).tocsr()
        K1 = api.get_operator(ctx, 1).tocsr()
        x0 = rng.standard_normal(K0.shape[0])
        x1 = rng.standard_normal(K1.shape[0])
        t0 = time.perf_counter()
        for _ in range(3):
            K0 @ x0
        t_sp0 = (time.perf_counter() - t0) / 3
        t0 = time.perf_counter()
        for _ in range(3):
            K1 @ x1
        t_sp1 = (time.perf_counter() - t0) / 3
        Kc = api.get_operator(ctx, mesh.nmg - 1).tocsc()
        t0 = time.perf_counter()
        lu = spla.splu(Kc)
        t_fac = time.perf_counter() - t0
        b = rng.standard_normal(Kc.shape[0])
        t0 = time.perf_counter()
        lu.solve(b)
        t_csol = time.perf_counter() - t0
    n_lvl1 = 6 if mesh.nmg > 2 else 0
    t_it = 7 * t_sp0 + n_lvl1 * t_sp1 + t_csol
    t_solve = t_asm + t_fac + n_cg * t_it
    sample = (f"oracle on one core: class kernels of {nsamp} sampled elements scaled to "
              f"{mesh.nel}; fine and level-2 SpMV x3 each, SuperLU factor+solve of the "
              f"coarse operator; combined with the GPU run's {n_cg} CG iterations "
              f"(7 fine + {n_lvl1} level-2 SpMVs + 1 coarse solve each); Galerkin RAP and "
              f"vector updates not counted")
    return dict(value=1.0 / t_solve, unit="solves/s", cores=1, kind="port", sample=sample,
                t_solve_s=t_solve, t_assembly_s=t_asm, t_cg_iter_s=t_it, t_coarse_factor_s=t_fac)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--form", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on exactly one GPU (the Stokes solve is single-device; "
              "multi-GPU sharding is a design note only). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from m3tet_sph_amd import mesh as M, fields as F, api

    t0 = time.perf_counter()
    mesh = M.build_config(args.config)
    inp = F.make_inputs(mesh, visc_contrast=1e4, Ra=1e5, seed_T=0, seed_plates=1)
    ctx = api.StokesContext(mesh)
    ctx.set_bc(inp["ifix"], inp["vfix"])
    log(f"setup {time.perf_counter() - t0:.1f}s config={args.config} nel={mesh.nel} "
        f"nU={3 * mesh.nnod} nP={mesh.nVnod} stats={ctx.stats()}")

    def solve(form):
        return ctx.flow3d(inp["Visc"], inp["Dens"], inp["phys"], P0=None, rtol_Pat=args.tol,
                          rtol_U_final=0.5 * args.tol, precision=form)

    # fixed denominator: pressure residual of the velocity solve at P = 0
    U, P, info = solve("fp64")
    rP_cold = info["norm_rP0"]

    def rel_res(U, P):
        rU, fU, rP = api.residual(ctx, U, P)
        return max(rU / fU, rP / rP_cold), rU / fU, rP / rP_cold

    def timed(form):
        for _ in range(args.warmup):
            solve(form)
        t0 = time.perf_counter()
        infos = []
        for _ in range(args.steps):
            U, P, info = solve(form)
            infos.append(info)
        return (time.perf_counter() - t0), U, P, infos

    plain = None
    if args.form != "fp64":
        dtp, Up, Pp, infp = timed("fp64")
        plain = dict(solves_per_s=round(args.steps / dtp, 4), ms_per_solve=round(1e3 * dtp / args.steps, 2),
                     rel_res=float(f"{rel_res(Up, Pp)[0]:.3e}"),
                     nitZ=infp[-1]["nitZ"], nitU=infp[-1]["nitU"], nitP=infp[-1]["nitP"])
        log(f"plain fp64: {plain}")
    dt, U, P, infos = timed(args.form)
    rr, rru, rrp = rel_res(U, P)
    ms = 1e3 * dt / args.steps
    err_vs_fp64 = None
    if plain is not None:
        err_vs_fp64 = dict(U=float(f"{np.linalg.norm(U - Up) / np.linalg.norm(Up):.3e}"),
                           P=float(f"{np.linalg.norm(P - Pp) / np.linalg.norm(Pp):.3e}"))
    n_cg = int(np.mean([i["nitU"] + i["nitZ"] for i in infos]))
    log(f"{args.steps} solves in {dt:.3f}s -> {ms:.1f} ms/solve; rel_res={rr:.2e} "
        f"(U {rru:.2e}, P {rrp:.2e}); last info {infos[-1]}")

    # roofline of the dominant kernel: fine-level fused Jacobi sweep
    api.profile(ctx, True)
    solve(args.form)
    pr = api.profile_read(ctx)
    api.profile(ctx, False)
    st = ctx.stats()
    nnzb, nnod = st["nnzb"][0], mesh.nnod
    if args.form == "fp64":
        nsw, tsw, bpl = pr["n_fp64"], pr["t_fp64"], sweep_bytes(nnzb, nnod, 8, 8)
    else:
        nsw, tsw, bpl = pr["n_mixed"], pr["t_mixed"], sweep_bytes(nnzb, nnod, 4, 4)
    avg = tsw / max(nsw, 1)
    achieved = bpl / avg / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel="k_smooth fine-level fused Jacobi sweep", launches=nsw,
                avg_us=round(avg * 1e6, 2), bytes_per_launch=int(bpl))
    log(f"sweep: {nsw} launches avg {avg * 1e6:.1f} us, {bpl / 1e6:.1f} MB -> {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(ctx, mesh, inp, n_cg)
        log(f"cpu baseline sample took {time.perf_counter() - t0:.1f}s: {cpu}")

    last = infos[-1]
    out = {
        "metric": "Stokes solves/s to rel.res 1e-6: mixed+learned vs plain FP64 MGPCG, 1 GPU",
        "value": round(1e3 / ms, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f64 outer, f32/bf16 V-cycle",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: spherical-shell Stokes, {mesh.nel} P2 tets, "
                               f"{3 * mesh.nnod} velocity dofs, {mesh.nVnod} pressure dofs, "
                               f"{mesh.nmg}-level MG, viscosity contrast 1e4",
                   "form": args.form, "rel_res_target": args.tol, "seeds": [0, 1, 2],
                   "Ra": 1e5},
        "rel_res": float(f"{rr:.3e}"),
        "iterations": {"nitP": last["nitP"], "nitU": last["nitU"], "nitZ": last["nitZ"]},
        "phases_ms": {k: round(1e3 * last[k], 2) for k in
                      ("t_assemble", "t_restrict", "t_factor", "t_pressure", "t_velocity")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (round(plain["ms_per_solve"] / ms, 3) if plain else None),
        "error_vs_fp64": err_vs_fp64,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
