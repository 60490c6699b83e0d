"""Headline benchmark: DMRG sweep time and H_eff TFLOP/s, mixed vs FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4|c2|small]

Workload (default "c4"): spin-1/2 Heisenberg XXX chain, L=128, open BC, J=1,
d=2, MPO bond D=5, chi_max=512, two-site DMRG, Lanczos Krylov cap 20,
seeded random initial MPS (splitmix64 seed 128, entries 0.1*U[-1,1)), bond
dimensions pre-sized to min(2^min(i,L-i), chi_max) as the reference does.

A step is one full sweep (left-to-right plus right-to-left).  For each of the
FP64 path and the mixed path (FP32 MFMA H_eff matvecs; FP64 environments,
Lanczos and SVD) the engine runs W warm-up sweeps from the seeded state, then
exactly K timed sweeps, one per engine call (the environment build before each
call is outside the sweep timer, as in the reference).  `value` is the mixed
path's mean seconds per sweep; the per-sweep times and their median are in
`modes`.  The
roofline entry is the H_eff matvec (three MFMA GEMM launches), timed with
HIP events on the engine's stream.  |dE| compares the two paths after the
same number of sweeps from the same state.

cpu_baseline: the reference CPU path (cpu/pdmrg tensordot H_eff + ARPACK
eigsh + numpy SVD, restated in oracle/pdmrg_serial.py) timed on this host for
a bounded sample of bonds of the same chain and state, scaled to seconds per
sweep.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dmrg-implementations_amd"))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense MFMA TFLOP/s

WORKLOADS = {
    "c4": dict(L=128, chi=512, krylov=20, seed=128),
    "c2": dict(L=64, chi=256, krylov=20, seed=64),
    "small": dict(L=32, chi=64, krylov=20, seed=32),
}


def heisenberg_mpo(L):
    """Same MPO as oracle.models.heisenberg_mpo (kept here so the GPU legs do
    not import the oracle)."""
    import numpy as np
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T.copy()
    Sz = np.diag([0.5, -0.5])
    I2 = np.eye(2)
    Ws = []
    for i in range(L):
        W = np.zeros((5, 2, 2, 5))
        blocks = []
        if i != L - 1:
            blocks += [(0, 0, I2), (0, 1, Sp), (0, 2, Sm), (0, 3, Sz)]
        if i != 0:
            blocks += [(1, 4, 0.5 * Sm), (2, 4, 0.5 * Sp), (3, 4, Sz), (4, 4, I2)]
        for w, wp, op in blocks:
            W[w, :, :, wp] = op.T
        Ws.append(W)
    return Ws


# Jacobi stop of the truncation split (max kept/discarded coupling relative to
# |rho|_F), the same in both modes: a looser 1e-12 in the mixed mode measured
# |dE|/|E| = 1.3e-10 against FP64 with no gain in split time.
SVD_TOL = {"fp64": 1e-14, "fp32": 1e-14}


def run_gpu(dmrg_amd, wl, precision, steps, warmup, svd_tol):
    Ws = heisenberg_mpo(wl["L"])
    eng = dmrg_amd.DMRG(wl["L"], 2, wl["chi"], 5, tol=0.0, krylov_max=wl["krylov"],
                        precision=precision, svd_tol=svd_tol)
    eng.set_mpo(Ws)
    eng.initialize_mps_random(seed=wl["seed"], scale=0.1)
    if warmup > 0:
        eng.run(warmup)
    eng.reset_stats()
    dmrg_amd.synchronize()
    # one sweep per call: each call rebuilds the environments (outside the
    # sweep timer, as in the reference) and times exactly one full sweep
    per_sweep = []
    t0 = time.perf_counter()
    for _ in range(steps):
        E = eng.run(1)
        per_sweep.append(eng.stats()["sweep_s_last"])
    dmrg_amd.synchronize()
    wall = time.perf_counter() - t0
    st = eng.stats()
    sweep_s = st["sweeps_s_total"] / max(steps, 1)
    heff_tf = st["heff_flop"] / (st["heff_ms"] * 1e-3) / 1e12 if st["heff_ms"] > 0 else 0.0
    res = dict(energy=E, sweep_s=sweep_s, wall_s=wall, env_build_s=st["env_build_s"],
               heff_tflops=heff_tf, heff_ms=st["heff_ms"], heff_flop=st["heff_flop"],
               matvecs=st["matvecs"], svd_ms=st["svd_ms"], lanczos_ms=st["lanczos_ms"],
               svd_calls=st["svd_calls"], svd_sweeps=st["svd_sweeps"],
               svd_warm_starts=st["svd_warm_starts"], svd_bip_sweeps=st["svd_bip_sweeps"],
               svd_oa_steps=st["svd_oa_steps"], sweep_s_each=per_sweep,
               sweep_s_median=sorted(per_sweep)[len(per_sweep) // 2] if per_sweep else 0.0,
               ms_per_matvec=st["heff_ms"] / max(st["matvecs"], 1))
    return eng, res


def cpu_baseline(eng, wl, nbonds):
    """Time the reference CPU path on `nbonds` bonds at the chain centre,
    starting from the engine's current MPS and environments."""
    import numpy as np
    from oracle import pdmrg_serial as ps
    L, chi = wl["L"], wl["chi"]
    Ws = [ps.to_pdmrg_mpo(W) for W in heisenberg_mpo(L)]
    mps = eng.get_mps()
    c0 = L // 2 - nbonds // 2
    L_envs = {c0: ps.to_pdmrg_env(eng.get_env(True, c0))}
    R_envs = {}
    for i in range(c0, c0 + nbonds):
        R_envs[i + 1] = ps.to_pdmrg_env(eng.get_env(False, i + 2))
    A = list(mps)
    t0 = time.perf_counter()
    for i in range(c0, c0 + nbonds):
        ps.bond_step(A, Ws, L_envs, R_envs, i, "right", chi, max_iter=wl["krylov"], tol=1e-10)
    t = time.perf_counter() - t0
    per_bond = t / nbonds
    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    return dict(value=per_bond * 2 * (L - 1), unit="s/sweep", cores=cores, kind="port",
                sample="%d centre bonds (L->R, chi=%d) of the same chain and state, %.2f s/bond, "
                       "scaled by the 2*(L-1) bond updates of a sweep" % (nbonds, chi, per_bond))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-bonds", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,fp32")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="before the runs, copy MIB MiB device-to-device 3 times (the known-size "
                         "copy kernel that scales the FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; the DMRG sweep of a single chain does not "
              "shard across GPUs here (see DESIGN.md).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import dmrg_amd
    wl = WORKLOADS[a.workload]
    if a.calib_copy > 0:
        import numpy as np
        n = a.calib_copy * (1 << 20) // 8
        src, dst = dmrg_amd.DeviceArray((n,), np.float64), dmrg_amd.DeviceArray((n,), np.float64)
        for _ in range(3):
            dmrg_amd.copy_device(dst, src)
        dmrg_amd.synchronize()
        src.free()
        dst.free()
    modes = a.modes.split(",")
    res, engines = {}, {}
    for m in modes:
        prec = dmrg_amd.PREC_FP64 if m == "fp64" else dmrg_amd.PREC_FP32
        print("[bench] %s: %d warm-up + %d timed sweeps, L=%d chi=%d" %
              (m, a.warmup, a.steps, wl["L"], wl["chi"]), file=sys.stderr, flush=True)
        eng, r = run_gpu(dmrg_amd, wl, prec, a.steps, a.warmup, SVD_TOL[m])
        r["svd_tol"] = SVD_TOL[m]
        res[m], engines[m] = r, eng
        print("[bench] %s: %s" % (m, json.dumps(r)), file=sys.stderr, flush=True)
    head = "fp32" if "fp32" in res else modes[0]
    cpu = None
    if not a.no_cpu:
        print("[bench] cpu baseline: %d bonds" % a.cpu_bonds, file=sys.stderr, flush=True)
        cpu = cpu_baseline(engines["fp64" if "fp64" in engines else head], wl, a.cpu_bonds)
    for e in engines.values():
        e.close()
    r = res[head]
    peak = PEAK["fp32" if head == "fp32" else "fp64"]
    out = {
        "metric": "sweep s + H_eff TFLOP/s at L=128 chi=512, mixed vs FP64; |dE| vs FP64",
        "value": r["sweep_s"],
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["sweep_s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 H_eff MFMA / f64" if head == "fp32" else "f64",
        "data": "synthetic (seeded random initial MPS)",
        "config": {"workload": "heisenberg_xxx L=%d chi_max=%d two-site D=5 d=2 krylov=%d seed=%d"
                               % (wl["L"], wl["chi"], wl["krylov"], wl["seed"]),
                   "L": wl["L"], "chi_max": wl["chi"], "krylov_max": wl["krylov"]},
        "roofline": {"bound": "mfma", "achieved": r["heff_tflops"], "peak": peak,
                     "unit": "TFLOP/s", "frac": r["heff_tflops"] / peak, "traffic": None},
        "cpu_baseline": cpu,
        "modes": res,
    }
    if "fp64" in res and "fp32" in res:
        e64, e32 = res["fp64"]["energy"], res["fp32"]["energy"]
        out["dE_abs"] = abs(e32 - e64)
        out["dE_rel"] = abs(e32 - e64) / abs(e64)
        out["speedup_mixed_vs_fp64"] = res["fp64"]["sweep_s"] / res["fp32"]["sweep_s"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
