"""Headline benchmark: DMRG sweep time and H_eff TFLOP/s, mixed vs FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4|c2|c3|c5|small]

Workload (default "c4"): spin-1/2 Heisenberg XXX chain, L=128, open BC, J=1,
d=2, MPO bond D=5, chi_max=512, two-site DMRG, Lanczos Krylov cap 20,
seeded random initial MPS (splitmix64 seed 128, entries 0.1*U[-1,1)), bond
dimensions pre-sized to min(2^min(i,L-i), chi_max) as the reference does.

A step is one full sweep (left-to-right plus right-to-left).  For each of the
FP64 path and the mixed paths (FP32 MFMA H_eff matvecs, and the same with split-BF16
products; FP64 environments, Lanczos and SVD) the engine runs W warm-up sweeps from the seeded state, then
exactly K timed sweeps, one per engine call (the environment build before each
call is outside the sweep timer, as in the reference).  The defaults, W=3 and
K=9, make the 12-sweep C4 run.  `value` is the mixed path's mean seconds per
timed sweep; every sweep's time (warm-up ones included), their mean over the
whole run and the median of the timed ones are in `modes`.  The roofline entry is the H_eff matvec (the first step fused with the
MPO-bond step, the split-K out-step slices and their sum; in the mixed modes
also the FP64/FP32 conversion of the two-site vector), timed with HIP events
on the engine's stream.  |dE| compares the two paths after the same number of sweeps from the
same state.

Other workloads (not the headline line): c2 (Heisenberg L=64 chi=256), c3
(complex Josephson array L=32 d=5 chi=256, FP64 only: the FP32 matvec is
real-only), c5 (64 uneven random-field chains: sequential plain FP64 solves
against dmrg_run_batch in FP64 and in the mixed mode; value = speedup).

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

# MI355X dense MFMA TFLOP/s; split-BF16 counts the BF16 peak over its six partial products
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 6, "fp32env": 157.3}
# mode -> (precision, environment precision); fp32env is the FP32 matvec with
# the environment updates on FP32 MFMA too (FP32 environments)
MODE = {"fp64": (0, 0), "fp32": (1, 0), "bf16x3": (2, 0), "fp32env": (1, 1)}

WORKLOADS = {
    # the headline: Heisenberg XXX L=128 chi=512, FP64 vs mixed
    "c4": dict(model="heisenberg", L=128, chi=512, krylov=20, seed=128),
    # parity-test configurations, measured on request
    "c2": dict(model="heisenberg", L=64, chi=256, krylov=20, seed=64),
    "c3": dict(model="josephson", L=32, chi=256, krylov=20, seed=32, d=5, D=4, dtype="c128",
               n_max=2, E_J=2.0, E_C=1.0, phi=3.141592653589793 / 3),
    "c5": dict(model="batch", chains=64, seed=5, krylov=20),
    "small": dict(model="heisenberg", L=32, chi=64, krylov=20, seed=32),
}


def _block(W, w, wp, op):
    """W[w, s, sp, wp] = op[sp, s] (op as a <bra|op|ket> matrix)."""
    W[w, :, :, wp] = op.T


def heisenberg_mpo(L, h=None):
    """Spin-1/2 XXX chain, D=5, with an optional field sum_i h_i Sz_i (same MPO
    as oracle.models.random_field_heisenberg_mpo; kept here so the GPU legs do
    not import the oracle)."""
    import numpy as np
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T.copy()
    Sz = np.diag([0.5, -0.5])
    I2 = np.eye(2)
    Ws = []
    for i in range(L):
        W = np.zeros((5, 2, 2, 5))
        if i != L - 1:
            for wp, op in ((0, I2), (1, Sp), (2, Sm), (3, Sz)):
                _block(W, 0, wp, op)
        if i != 0:
            for w, op in ((1, 0.5 * Sm), (2, 0.5 * Sp), (3, Sz), (4, I2)):
                _block(W, w, 4, op)
        if h is not None and h[i] != 0.0:
            _block(W, 0, 4, h[i] * Sz)
        Ws.append(W)
    return Ws


def josephson_mpo(L, n_max, E_J, E_C, phi):
    """Josephson-junction array, complex, D=4 (same MPO as
    oracle.models.josephson_mpo): E_C n^2 on site, -E_J/2 e^{i phi} on
    e^{i phi_i} e^{-i phi_{i+1}} plus its conjugate."""
    import numpy as np
    d = 2 * n_max + 1
    q = np.arange(d) - n_max
    onsite = np.diag(E_C * q * q).astype(complex)
    up = np.zeros((d, d), complex)
    for k in range(d - 1):
        up[k + 1, k] = 1.0
    down = up.conj().T
    a = -0.5 * E_J * np.exp(1j * phi)
    eye = np.eye(d, dtype=complex)
    Ws = []
    for i in range(L):
        W = np.zeros((4, d, d, 4), complex)
        if i != L - 1:
            _block(W, 0, 0, eye)
            _block(W, 0, 1, a * up)
            _block(W, 0, 2, np.conj(a) * down)
        _block(W, 0, 3, onsite)
        if i != 0:
            _block(W, 1, 3, down)
            _block(W, 2, 3, up)
            _block(W, 3, 3, eye)
        Ws.append(W)
    return Ws


def workload_mpo(wl):
    if wl["model"] == "josephson":
        return josephson_mpo(wl["L"], wl["n_max"], wl["E_J"], wl["E_C"], wl["phi"])
    return heisenberg_mpo(wl["L"])


# Jacobi stop of the truncation split (max kept/discarded coupling relative to
# |rho|_F): the engine's defaults, 1e-14 in FP64 and 1e-13 in the mixed modes
# (whose two-site vectors carry FP32 rounding; measured at C4: mixed split
# 4.1 s instead of 5.2 s per 3 sweeps, |dE|/|E| 8e-12 against FP64; 1e-12
# gave 1.3e-10).
SVD_TOL = {"fp64": 1e-14, "fp32": 1e-13, "bf16x3": 1e-13, "fp32env": 1e-13}


def run_gpu(dmrg_amd, wl, precision, steps, warmup, svd_tol, env_precision=0, rsvd=-1):
    Ws = workload_mpo(wl)
    eng = dmrg_amd.DMRG(wl["L"], wl.get("d", 2), wl["chi"], wl.get("D", 5),
                        dtype=wl.get("dtype", "f64"), tol=0.0, krylov_max=wl["krylov"],
                        precision=precision, svd_tol=svd_tol, svd_tol_mixed=svd_tol, profile=1,
                        env_precision=env_precision, svd_rsvd=1 if rsvd >= 0 else 0,
                        rsvd_power=max(rsvd, 0), autotune=1)
    eng.set_mpo(Ws)
    eng.initialize_mps_random(seed=wl["seed"], scale=0.1)
    warm_sweeps = []
    for _ in range(warmup):  # one sweep per call, timed apart (reported, not in `value`)
        eng.run(1)
        warm_sweeps.append(eng.stats()["sweep_s_last"])
    eng.reset_stats()
    dmrg_amd.synchronize()
    # one sweep per call: each call rebuilds the environments (outside the
    # sweep timer, as in the reference) and times exactly one full sweep
    per_sweep = []
    E = eng.get_energy()
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
               warmup_sweep_s=warm_sweeps,
               sweep_s_mean_all=(sum(warm_sweeps) + sum(per_sweep)) / max(len(warm_sweeps) + len(per_sweep), 1),
               ms_per_matvec=st["heff_ms"] / max(st["matvecs"], 1),
               env_ms=st["env_ms"], env_flop=st["env_flop"],
               env_tflops=st["env_flop"] / (st["env_ms"] * 1e-3) / 1e12 if st["env_ms"] > 0 else 0.0,
               trunc_err_last_sweep=st["trunc_sum_last"], max_discarded=st["max_discarded"],
               env_precision="fp32" if (precision != 0 and eng.opts.env_precision) else "fp64")
    return eng, res


def sweep_bond_equivalents(wl, c0, nbonds):
    """A sweep's 2(L-1) bond updates in units of the mean of the measured
    bonds c0 .. c0+nbonds-1, weighting each bond by its H_eff flops
    F_heff2(cL, cR) at the pre-sized bond dimensions min(d^min(i, L-i), chi)
    (the H_eff dominates the CPU bond step)."""
    L, chi = wl["L"], wl["chi"]
    d, D = wl.get("d", 2), wl.get("D", 5)
    bonds = [min(d ** min(i, L - i), chi) for i in range(L + 1)]
    f = lambda cl, cr: 2.0 * D * d * d * cl * cr * (cl + cr) + 2.0 * cl * cr * (D * d * d) ** 2
    w = [f(bonds[i], bonds[i + 2]) for i in range(L - 1)]
    w_meas = sum(w[c0:c0 + nbonds]) / nbonds
    return 2 * sum(w) / w_meas


def cpu_baseline(eng, wl, nbonds):
    """Time the reference CPU path on `nbonds` bonds at the chain centre,
    starting from the engine's current MPS and environments."""
    import numpy as np
    from oracle import pdmrg_serial as ps
    L, chi = wl["L"], wl["chi"]
    Ws = [ps.to_pdmrg_mpo(W) for W in workload_mpo(wl)]
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
    bond_equiv = sweep_bond_equivalents(wl, c0, nbonds)
    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    return dict(value=per_bond * bond_equiv, unit="s/sweep", cores=cores, kind="port",
                sample="%d centre bonds (L->R, chi=%d) of the same chain and state, %.2f s/bond, "
                       "scaled to a sweep by the H_eff flops of its 2*(L-1) bond updates "
                       "(%.1f centre-bond equivalents)" % (nbonds, chi, per_bond, bond_equiv))


def heff_traffic(workload, mode):
    """HBM bytes per H_eff matvec at the C4 centre bond in the mixed mode, from
    the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes (copy-calibrated);
    None for other workloads and modes."""
    path = os.path.join(ROOT, "profiles", "c4_heff_traffic.json")
    if workload != "c4" or mode != "fp32" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["bytes_per_matvec"]


METRIC_C4 = "sweep s + H_eff TFLOP/s at L=128 \u03c7=512, mixed vs FP64; |\u0394E| vs FP64"


def batch_chains(wl):
    """64 random-field Heisenberg chains: L in {16,24,32,48,64}, chi_max in
    {32,64,128,256}, disorder W in {0.5,1,2,5}, h_i ~ U[-W, W] (seeded)."""
    import numpy as np
    rng = np.random.default_rng(wl["seed"])
    out = []
    for _ in range(wl["chains"]):
        L = int(rng.choice([16, 24, 32, 48, 64]))
        chi = int(rng.choice([32, 64, 128, 256]))
        W = float(rng.choice([0.5, 1.0, 2.0, 5.0]))
        out.append((L, chi, W, rng.uniform(-W, W, L)))
    return out


def run_batch_bench(dmrg_amd, wl, a):
    """Sequential plain FP64 solves of every chain, then the same chains run
    together by dmrg_run_batch (FP64, and FP32-matvec mode); each batched
    energy is compared with its chain's own FP64 solve."""
    import numpy as np
    chains = batch_chains(wl)
    sweeps = a.batch_sweeps
    ref, t_seq = [], 0.0
    for i, (L, chi, W, h) in enumerate(chains):
        e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=wl["krylov"], autotune=1)
        e.set_mpo(heisenberg_mpo(L, h))
        e.initialize_mps_random(seed=1000 + i, scale=0.1)
        dmrg_amd.synchronize()
        t0 = time.perf_counter()
        ref.append(e.run(sweeps))
        dmrg_amd.synchronize()
        t_seq += time.perf_counter() - t0
        e.close()
    print("[bench] c5 sequential fp64: %.2f s for %d chains" % (t_seq, len(chains)),
          file=sys.stderr, flush=True)
    out_modes = {}
    for m, prec in (("fp64", dmrg_amd.PREC_FP64), ("fp32", dmrg_amd.PREC_FP32)):
        engines = []
        for i, (L, chi, W, h) in enumerate(chains):
            e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=wl["krylov"], precision=prec,
                              graph=a.graph, autotune=1)
            e.set_mpo(heisenberg_mpo(L, h))
            e.initialize_mps_random(seed=1000 + i, scale=0.1)
            engines.append(e)
        dmrg_amd.synchronize()
        t0 = time.perf_counter()
        E, T = dmrg_amd.run_batch(engines, sweeps, threads=a.batch_threads)
        dmrg_amd.synchronize()
        t_b = time.perf_counter() - t0
        for e in engines:
            e.close()
        rel = [abs(E[i] - ref[i]) / abs(ref[i]) for i in range(len(chains))]
        out_modes[m] = dict(batch_s=t_b, speedup_vs_sequential=t_seq / t_b, max_dE_rel=max(rel),
                            median_dE_rel=float(np.median(rel)))
        print("[bench] c5 batch %s: %s" % (m, json.dumps(out_modes[m])), file=sys.stderr, flush=True)
    head = out_modes["fp32"]
    print(json.dumps({
        "metric": "64-chain uneven batch: sequential plain FP64 time / batched time (mixed); "
                  "max |\u0394E|/|E| vs each chain's own FP64 solve",
        "value": head["speedup_vs_sequential"], "unit": "x", "n_gpus": 1, "steps": sweeps,
        "warmup": 0, "ms_per_step": head["batch_s"] * 1e3 / sweeps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 H_eff MFMA / f64",
        "data": "synthetic (seeded random-field chains and initial MPS)",
        "config": {"workload": "64 random-field Heisenberg chains, L in {16,24,32,48,64}, chi_max in "
                               "{32,64,128,256}, W in {0.5,1,2,5}, seed %d, %d sweeps each, %d host "
                               "threads, Lanczos graphs %s" % (wl["seed"], sweeps, a.batch_threads,
                                                               "on" if a.graph else "off"),
                   "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default")},
        "sequential_fp64_s": t_seq, "modes": out_modes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # C4 is a 12-sweep run: 3 sweeps from the random start (untimed in `value`),
    # then 9 timed sweeps
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-bonds", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,fp32,bf16x3")
    ap.add_argument("--batch-sweeps", type=int, default=6, help="c5: sweeps per chain")
    ap.add_argument("--batch-threads", type=int, default=8, help="c5: host threads (chains in flight)")
    ap.add_argument("--rsvd", type=int, default=-1, metavar="POWER",
                    help="randomised split with POWER subspace iterations on the bonds where the "
                         "reference uses it (-1: exact split everywhere)")
    ap.add_argument("--graph", type=int, default=1,
                    help="c5: batched chains replay their Lanczos stretches as HIP graphs (1) "
                         "or launch every kernel (0)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="set GPU_MAX_HW_QUEUES (hardware queues of this process, at most 32) "
                         "before the runtime starts; 0 leaves the environment as it is.  The "
                         "batch runner's chains run on their own streams, which share these "
                         "queues")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="before the runs, copy MIB MiB device-to-device 3 times (the known-size "
                         "copy kernel that scales the FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args()
    bad = [m for m in a.modes.split(",") if m not in PEAK]
    if bad:
        ap.error("--modes: unknown mode(s) %s; choose from %s" % (",".join(bad), ",".join(PEAK)))
    if a.steps < 1 or a.warmup < 0:
        ap.error("--steps must be >= 1 and --warmup >= 0")
    if a.gpus != 1:
        print("bench.py: measures one GPU only; the sweep of a single chain is sequential over its "
              "bonds and is not sharded across GPUs.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.hw_queues:
        if not 1 <= a.hw_queues <= 32:
            ap.error("--hw-queues must be in [1, 32]")
        os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)
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
    if wl["model"] == "batch":
        run_batch_bench(dmrg_amd, wl, a)
        return
    modes = a.modes.split(",")
    if wl.get("dtype") == "c128":
        modes = [m for m in modes if m == "fp64"] or ["fp64"]  # the FP32 matvec is real-only
    res, engines = {}, {}
    for m in modes:
        prec, envp = MODE[m]
        print("[bench] %s: %d warm-up + %d timed sweeps, L=%d chi=%d" %
              (m, a.warmup, a.steps, wl["L"], wl["chi"]), file=sys.stderr, flush=True)
        eng, r = run_gpu(dmrg_amd, wl, prec, a.steps, a.warmup, SVD_TOL[m], envp, a.rsvd)
        r["rsvd_calls"] = eng.stats()["rsvd_calls"]
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
    peak = PEAK[head]
    if a.workload == "c4":
        metric = METRIC_C4
    else:
        metric = "sweep s + H_eff TFLOP/s at %s L=%d \u03c7=%d" % (wl["model"], wl["L"], wl["chi"])
    if wl["model"] == "josephson":
        desc = ("josephson_array L=%d d=%d D=4 chi_max=%d complex128 E_J=%g E_C=%g flux=pi/3 "
                "two-site krylov=%d seed=%d" % (wl["L"], wl["d"], wl["chi"], wl["E_J"], wl["E_C"],
                                                 wl["krylov"], wl["seed"]))
    else:
        desc = ("heisenberg_xxx L=%d chi_max=%d two-site D=5 d=2 krylov=%d seed=%d"
                % (wl["L"], wl["chi"], wl["krylov"], wl["seed"]))
    dtype = "f32 H_eff MFMA / f64" if head == "fp32" else ("c128" if wl.get("dtype") == "c128" else "f64")
    out = {
        "metric": metric,
        "value": r["sweep_s"],
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["sweep_s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded random initial MPS)",
        "config": {"workload": desc, "L": wl["L"], "chi_max": wl["chi"], "krylov_max": wl["krylov"]},
        "roofline": {"bound": "mfma", "achieved": r["heff_tflops"], "peak": peak,
                     "unit": "TFLOP/s", "frac": r["heff_tflops"] / peak,
                     "traffic": heff_traffic(a.workload, head)},
        "cpu_baseline": cpu,
        "modes": res,
    }
    if "fp64" in res and "fp32" in res:
        e64, e32 = res["fp64"]["energy"], res["fp32"]["energy"]
        out["dE_abs"] = abs(e32 - e64)
        out["dE_rel"] = abs(e32 - e64) / abs(e64)
        out["speedup_mixed_vs_fp64"] = res["fp64"]["sweep_s"] / res["fp32"]["sweep_s"]
    # truncation error (discarded weight summed over the last sweep's splits)
    # of every mode, next to the FP64 run's
    out["trunc_err_last_sweep"] = {m: res[m]["trunc_err_last_sweep"] for m in res}
    for m in res:
        if m != "fp64" and "fp64" in res:
            res[m]["dE_rel_vs_fp64"] = abs(res[m]["energy"] - res["fp64"]["energy"]) / abs(res["fp64"]["energy"])
            res[m]["speedup_vs_fp64"] = res["fp64"]["sweep_s"] / res[m]["sweep_s"]
            t64 = res["fp64"]["trunc_err_last_sweep"]
            res[m]["trunc_err_rel_vs_fp64"] = (abs(res[m]["trunc_err_last_sweep"] - t64) / t64
                                               if t64 > 0 else None)
        res[m]["heff_frac_of_peak"] = res[m]["heff_tflops"] / PEAK[m]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
