"""S(Q,w) benchmark on one MI355X.

Workload (the configuration the metric is quoted on): cubic ABO3 perovskite,
40^3 cells = 320,000 atoms in 3 species, 512 MD-like frames at dt = 10 fs
(64 commensurate plane-wave phonons, rms 0.06 A, plus 0.01 A noise), a
per-species type-1 NUFFT to 256 x 256 x 32 modes at eps 1e-7, species combine
with neutron weights, one Hann window over the 512 frames (n_int = 512,
n_frame = 1), time FFT and |.|^2 -> S(Q,w) for m = 0..255 plus S(Q).

A step is one full pass: 512 frames x 3 species NUFFTs + combine + reduction,
inputs resident in HBM.  value = frames/s of the FP32 form (FP32 spread /
FFT / deconvolution, FP64 fold and FP64 final sums).  The plain FP64 form
runs beside it on the same trajectory and is the reference for the FP32
form's error; the CPU leg times the host FINUFFT-method restatement
(oracle/) on a bounded sample of the same frames.

Beside the step the line carries: the plain FP64 form and the FP32 form's
error against it; the 128^3 precision sweep (FP32 / FP64 time and error at
10,000 modes per eps); a serving record (Poisson load, p50/p99 latency);
and, in the CPU leg, the host baseline (all cores and one core) and the
oracle checks of S(Q,w) and of two served maps against the FP64 direct DFT.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--frames T] [--eps E]
                       [--no-fp64] [--no-cpu] [--no-c3] [--no-serve] [--cpu-frames F]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=512)
    ap.add_argument("--ncell", type=int, default=40)
    ap.add_argument("--modes", type=str, default="256,256,32")
    ap.add_argument("--eps", type=float, default=1e-7)
    ap.add_argument("--check-q", type=int, default=1,
                    help="check S(Q,w) against the FP64 direct DFT at --check-sets seeded Q (0: off)")
    ap.add_argument("--check-sets", type=str, default="8,8,4",
                    help="sizes of the seeded random k1, k2, k3 sets whose product is checked")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=20)
    ap.add_argument("--cpu-frames-1core", type=int, default=2)
    ap.add_argument("--no-c3", action="store_true", help="skip the 128^3 precision-sweep record")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving record")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default spread tile instead of timing the shapes first")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the record of the cut / I(Q,t) / one-phonon / DOS paths")
    ap.add_argument("--serve-rates", type=str, default="10,50,150")
    ap.add_argument("--serve-requests", type=int, default=40, help="requests per rate of the serving record")
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--workload", choices=("sqw", "batch", "serve", "sql"), default="sqw",
                    help="sqw: the S(Q,w) trajectory benchmark (default line); batch: 64 uneven jobs "
                         "in one launch per phase; serve: the request loop")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 launches) and exit: "
                         "the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    a.check_sets = tuple(int(v) for v in a.check_sets.split(","))
    a.serve_rates = tuple(float(v) for v in a.serve_rates.split(","))
    return a


def _flat_index(N, k1s, k2s, k3s):
    """CMCL flat indices of a product set of modes, in its [k3][k2][k1] order."""
    K3, K2, K1 = np.meshgrid(k3s, k2s, k1s, indexing="ij")
    return ((K3 + N[2] // 2) * N[1] * N[0] + (K2 + N[1] // 2) * N[0] + (K1 + N[0] // 2)).ravel()


def _mode_set(rng, N, sizes):
    return [np.sort(rng.choice(np.arange(-(n // 2), (n - 1) // 2 + 1), k, replace=False))
            for n, k in zip(N, sizes)]


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N > 1 is not supported "
                                   "(the multi-GPU sharding is a design only)"}))
        sys.exit(2)
    import torch

    if a.copy_mib:
        copy_calibration(a.copy_mib)
        return
    if a.workload == "batch":
        batch_bench(a)
        return
    if a.workload == "serve":
        serve_bench(a)
        return
    if a.workload == "sql":
        sql_bench(a)
        return
    from mp_tools_amd import synth
    from mp_tools_amd.sqw import SqwEngine

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    N = tuple(int(v) for v in a.modes.split(","))
    NQ = math.prod(N)
    T, ncell = a.frames, a.ncell
    t0 = time.time()
    pos, counts = synth.phonon_trajectory(ncell, T, dt_ps=0.01, n_phonons=64, rms=0.06,
                                          noise=0.01, seed=a.seed, device=dev)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    M = int(counts.sum())
    nat = float(M)
    w = 0.1 * synth.B_NEUTRON_FM                      # b_coh * 0.1 (mp_sqom56.f90:473)
    geo = synth.cell_geometry(ncell)
    n_f = T // 2
    dt = 0.01

    def make(precision):
        return SqwEngine(3, N, counts, eps=a.eps, precision=precision, weights=w,
                         tpe=geo["tpe"], batch_frames=a.batch)

    def step(eng, out):
        amp = eng.amplitudes(pos, out=out["amp"])
        eng.reduce(amp, n_int=T, win_kind=0, dt=dt, nat=nat, m_lo=0, n_f=n_f,
                   out=out["S"], out_tot=out["St"])

    def run(precision, steps, warmup):
        eng = make(precision)
        cdt = torch.complex128 if precision == "fp64" else torch.complex64
        out = dict(amp=torch.empty((T, NQ), dtype=cdt, device=dev),
                   S=torch.empty((n_f, NQ), dtype=torch.float64, device=dev),
                   St=torch.empty(NQ, dtype=torch.float64, device=dev))
        for _ in range(warmup):
            step(eng, out)
        torch.cuda.synchronize()
        eng.set_timing(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            step(eng, out)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        tm = eng.timing()
        return eng, out, el / steps, tm

    # spread tile shapes timed on the first frames and cached per geometry
    # (mpsq_autotune_tiles), as a user would before a long run; not timed
    tuned = {}
    if not a.no_autotune:
        from mp_tools_amd.sqw import autotune_tiles
        for prec in ("fp32",) + (() if a.no_fp64 else ("fp64",)):
            t0 = time.time()
            by, bz, msf = autotune_tiles(pos, 3, N, counts, eps=a.eps, precision=prec, weights=w, tpe=geo["tpe"])
            tuned[prec] = {"tile": [by, bz], "ms_per_frame": msf, "s": time.time() - t0}
    eng32, out32, sec32, tm32 = run("fp32", a.steps, a.warmup)
    frames_s = T / sec32
    info32 = eng32.info()
    res = {
        "metric": "S(Q,w) frames/s (FP32 form, rel.err vs FP64)",
        "value": frames_s,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sec32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded perovskite phonon trajectory generated on the GPU)",
        "config": {"workload": f"perovskite {ncell}^3 cells ({M} atoms, 3 species) x {T} frames, "
                               f"per-species type-1 NUFFT to {N[0]}x{N[1]}x{N[2]} modes eps {a.eps:g}, "
                               f"neutron-weighted combine, Hann time FFT n_int={T}, S(Q,w) m=0..{n_f - 1}",
                   "frames": T, "atoms": M, "species": 3, "modes": list(N), "eps": a.eps,
                   "n_int": T, "n_f": n_f, "seed": a.seed,
                   "fine_grid": [info32["n1"], info32["n2"], info32["n3"]], "kernel_width": info32["w"],
                   "spread_tile": [info32["by"], info32["bz"]]},
        "nu_pts_per_s": T * M / sec32,
        "autotune": tuned,
        "trajectory_gen_s": gen_s,
    }
    # stage times (HIP events on the engine's stream over the timed steps)
    stages = {}
    for k, (ms, n) in tm32.items():
        if n:
            stages[k] = {"ms_per_step": ms / a.steps, "launches_per_step": n / a.steps,
                         "avg_ms_per_launch": ms / n}
    res["stages_fp32"] = stages
    # roofline of the dominant kernel (per-kernel algorithmic bytes)
    n_fine = (info32["n1"], info32["n2"], info32["n3"])
    s_c = 8
    # scalar species weights: a frame is one transform over all atoms with
    # strengths w_s c_j, so every pass after the spread moves one plane per frame;
    # q_center 0: the strengths are real, and the x and y passes carry the
    # half spectrum (N1/2 + 1 x frequencies, N2 + 1 y frequencies) whose
    # mirrors the final pass writes
    H1, N2o = N[0] // 2 + 1, N[1] + 1
    alg = {
        "spread_xfft": T * (M * (12 + s_c) + s_c * n_fine[2] * n_fine[1] * H1),
        "colfft": T * s_c * (n_fine[2] * n_fine[1] * H1 + n_fine[2] * N2o * H1),
        "final_colfft": T * s_c * (n_fine[2] * N2o * H1 + NQ),
        "reduce": T * NQ * s_c + NQ * (n_f + 1) * 8,
        "sort": T * M * (4 + 2 * 32 + 2 * s_c),
        "stage": T * M * (16 + 32 + 4 + s_c),
    }
    dom = max(stages, key=lambda k: stages[k]["ms_per_step"])
    launches = stages[dom]["launches_per_step"]
    per_launch_bytes = alg.get(dom, 0) / launches
    achieved = per_launch_bytes / (stages[dom]["avg_ms_per_launch"] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_c2_fp32.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            if tj.get("eps") in (None, a.eps) and dom in tj:
                traffic = tj[dom]
        except Exception:
            traffic = None
    res["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                       "algorithmic_bytes_per_launch": per_launch_bytes}
    if dom == "spread_xfft":
        # the same launch priced with the complex form's output (all N1 x
        # frequencies), the byte model of a transform of complex strengths
        full = T * (M * (12 + s_c) + s_c * n_fine[2] * n_fine[1] * N[0]) / launches
        res["roofline"]["frac_full_spectrum_bytes"] = (
            full / (stages[dom]["avg_ms_per_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    # the spread kernel is bound by VALU issue, not HBM: its measured share of
    # the SIMDs' VALU issue slots (SQ_INSTS_VALU x 4 cycles / (SIMDs x cycles))
    cfile = os.path.join(ROOT, "profiles", "c2_fp32_instruction_counters.txt")
    if dom == "spread_xfft" and os.path.exists(cfile):
        for line in open(cfile):
            if line.startswith("k_spread_narrow") and line.rstrip().endswith("%"):
                res["roofline"]["valu_issue_frac"] = float(line.split()[-1].rstrip("%")) / 100.0
                res["roofline"]["valu_issue_source"] = "profiles/c2_fp32_instruction_counters.txt"
                break
    for k in stages:
        if k in alg:
            stages[k]["alg_GBps"] = alg[k] / (stages[k]["ms_per_step"] * 1e-3) / 1e9
    # the whole step on its own algorithmic bytes (every kernel of the step)
    step_bytes = sum(alg.values())
    res["step_alg_bytes"] = step_bytes
    res["step_alg_GBps"] = step_bytes / sec32 / 1e9
    res["step_alg_frac"] = res["step_alg_GBps"] / HBM_PEAK_GBS
    sf = sum(alg[k] for k in ("spread_xfft", "colfft", "final_colfft"))
    st = sum(stages[k]["ms_per_step"] for k in ("spread_xfft", "colfft", "final_colfft") if k in stages)
    res["spread_fft_alg_frac"] = sf / (st * 1e-3) / 1e9 / HBM_PEAK_GBS

    # 256 seeded random Q (a product set 8 x 8 x 4) for the direct-DFT check
    qsets = _mode_set(np.random.default_rng(a.seed + 1), N, a.check_sets)
    qidx = torch.tensor(_flat_index(N, *qsets), device=dev)
    S_sel = {"fp32": (out32["S"][:, qidx].cpu().numpy(), out32["St"][qidx].cpu().numpy())}

    # plain FP64 form beside it, and the FP32 form's error against it
    if not a.no_fp64:
        eng64, out64, sec64, tm64 = run("fp64", max(1, min(a.steps, 2)), 1)
        S32, S64 = out32["S"], out64["S"]
        St32, St64 = out32["St"], out64["St"]
        err_s = (torch.linalg.vector_norm(S32 - S64) / torch.linalg.vector_norm(S64)).item()
        err_t = (torch.linalg.vector_norm(St32 - St64) / torch.linalg.vector_norm(St64)).item()
        # the same without Bragg pixels (k_d multiple of ncell on the full grid)
        ks = [torch.arange(-(v // 2), (v - 1) // 2 + 1, device=dev) for v in N]
        K3, K2, K1 = torch.meshgrid(ks[2], ks[1], ks[0], indexing="ij")
        bragg = ((K1 % ncell == 0) & (K2 % ncell == 0) & (K3 % ncell == 0)).reshape(-1)
        keep = ~bragg
        err_nb = (torch.linalg.vector_norm((S32 - S64)[:, keep]) /
                  torch.linalg.vector_norm(S64[:, keep])).item()
        res["fp64_frames_per_s"] = T / sec64
        res["fp32_over_fp64_speedup"] = sec64 / sec32
        res["rel_l2_S_fp32_vs_fp64"] = err_s
        res["rel_l2_S_fp32_vs_fp64_no_bragg"] = err_nb
        res["rel_l2_Stot_fp32_vs_fp64"] = err_t
        res["stages_fp64"] = {k: {"ms_per_step": ms / max(1, min(a.steps, 2))}
                              for k, (ms, n) in tm64.items() if n}
        S_sel["fp64"] = (S64[:, qidx].cpu().numpy(), St64[qidx].cpu().numpy())
        del out64, eng64
        torch.cuda.empty_cache()
    if not a.no_next:
        res["next_rows"] = next_rows_record(a, dev, eng32, out32["amp"], pos, counts, nat, dt)
    del eng32
    amp_head = out32["amp"][:max(1, a.cpu_frames)].cpu().numpy()
    del out32
    torch.cuda.empty_cache()

    # precision sweep on the 128^3 static supercell (C3)
    c3 = None
    if not a.no_c3:
        c3 = c3_record(a, dev)
        res["c3_precision_sweep"] = c3["record"]
    # serving: Poisson load against resident trajectories (C5)
    srv = None
    if not a.no_serve:
        srv = serve_record(a, dev, pos, counts, c3["pos"] if c3 else None)
        res["serving"] = srv["record"]

    # CPU leg: the oracle's host FINUFFT-method restatement timed on a bounded
    # sample, and the oracle's FP64 direct-DFT checks of the GPU results
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_leg(a, pos, counts, w, N, amp_head, geo)
        if a.check_q:
            res["oracle_check"] = direct_dft_check(a, pos, counts, w, qsets, S_sel, geo, n_f, dt, nat)
        if srv is not None:
            res["serving"]["served_map_check"] = served_map_check(srv)
    print(json.dumps(res))


def _timed(fn, reps=3):
    """Median wall time (ms) of fn() on the current stream, synchronised."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))


def next_rows_record(a, dev, eng, amp, pos, counts, nat, dt):
    """Times of the paths either side of the S(Q,w) map on the same workload:
    an E(Q) cut, E = const maps over sliding windows, and I(Q,t) (map at 4 lags by direct sums and at every lag by
    FFT correlation, cut at every lag) from the C2
    amplitudes, the one-phonon amplitudes of the C2 lattice (5 sites x
    40^3 cells, 64 frames) and the mp_dos spectra of its velocities (central
    differences of the positions, all frames).  Algorithmic bytes: amplitudes / positions /
    velocities read once plus the outputs written."""
    import torch
    from mp_tools_amd import synth
    from mp_tools_amd import sqw as S
    T, NQ = amp.shape
    rec = {}
    g = S.qplane((40, 40, 40), 2, (1, 0, 0), (0, 1, 0))
    # a Q line across the plane's pixels, mapped onto the first 256 x 256 modes (kz = 0 plane)
    N = (256, 256, 32)
    cut = S.qcut(dict(g, n_qx=N[0], n_qy=N[1]), (-0.9, -0.6, 0.0), (0.95, 0.7, 0.0), (128, 128, 128))
    qi = torch.tensor(cut["qidx"] + N[0] * N[1] * (N[2] // 2), dtype=torch.int32, device=dev)
    nq = int(qi.numel())
    n_int, n_fr = S.reference_n_int(T, 0, sqw=True)
    ms = _timed(lambda: eng.reduce(amp, n_int=0, odd_n_int=True, nat=nat, dt=dt, qidx=qi))
    rec["eq_cut"] = {"pixels": nq, "n_int": n_int, "frequencies": n_fr, "ms": ms,
                     "alg_GBps": (T * nq * 8 + n_fr * nq * 8) / (ms * 1e-3) / 1e9}
    # E = const maps with the reference's default window (n_int = nfile / 2,
    # odd: 256 sliding windows), 17 energies around 0 over all modes
    ms = _timed(lambda: eng.reduce(amp, n_int=0, odd_n_int=True, m_lo=-8, n_f=17, nat=nat, dt=dt), reps=2)
    rec["econst_map_sliding"] = {"modes": NQ, "n_int": n_int, "windows": T - n_int + 1, "energies": 17, "ms": ms,
                                 "alg_GBps": (T * NQ * 8 + 18 * NQ * 8) / (ms * 1e-3) / 1e9}
    n_i, n_lag = S.reference_n_int(T, 0, sqw=False)
    ms = _timed(lambda: eng.reduce(amp, n_int=0, odd_n_int=True, kind=2, nat=nat, qidx=qi))
    rec["iqt_cut"] = {"pixels": nq, "n_int": n_i, "lags": n_lag, "ms": ms}
    ms = _timed(lambda: eng.reduce(amp, n_int=0, odd_n_int=True, kind=1, m_lo=0, n_f=4, nat=nat), reps=2)
    rec["iqt_map_4_lags"] = {"modes": NQ, "n_int": n_i, "ms": ms,
                             "alg_GBps": (T * NQ * 8 + 5 * NQ * 8) / (ms * 1e-3) / 1e9}
    ms = _timed(lambda: eng.reduce(amp, n_int=0, odd_n_int=True, kind=1, nat=nat), reps=2)
    rec["iqt_map_all_lags"] = {"modes": NQ, "n_int": n_i, "lags": n_lag, "ms": ms, "method": "FFT correlation",
                               "alg_GBps": (T * NQ * 8 + (n_lag + 1) * NQ * 8) / (ms * 1e-3) / 1e9}
    torch.cuda.empty_cache()
    # one-phonon amplitudes of the lattice (CELL layout: 5 site blocks)
    R, _ = synth.perovskite_lattice(40)
    cells = np.floor(R + 20.0 + 1e-9).astype(np.int32)
    at_ind = np.zeros((R.shape[0], 4), np.int32)
    at_ind[:, :3] = cells + 1
    at_ind[:, 3] = np.repeat(np.arange(1, 6), 40 ** 3)
    dind = torch.tensor(at_ind, device=dev)
    nfr = min(T, 64)
    ms = _timed(lambda: S.oneph_amplitudes(pos[:nfr], dind, synth.SITES, (40, 40, 40), 2, 1, (0.0, 0.0),
                                           weights=np.array([5.07, 7.16, 5.80, 5.80, 5.80]) * 0.1))
    npx = (2 * 40) ** 2
    rec["one_phonon"] = {"frames": nfr, "atoms": int(R.shape[0]), "pixels": npx, "ms": ms,
                         "ms_per_frame": ms / nfr,
                         "alg_GBps": nfr * (R.shape[0] * 16 + npx * 8) / (ms * 1e-3) / 1e9}
    # mp_dos spectra of the velocities (central differences, one-sided at the ends)
    vel = torch.zeros_like(pos)
    vel[1:-1, :, :3] = (pos[2:, :, :3] - pos[:-2, :, :3]) / (2 * dt)
    vel[0, :, :3] = (pos[1, :, :3] - pos[0, :, :3]) / dt
    vel[-1, :, :3] = (pos[-1, :, :3] - pos[-2, :, :3]) / dt
    ntype = 3
    ind3 = torch.tensor(np.repeat(np.arange(1, 4), counts).astype(np.int32)[:, None] * np.array([[0, 0, 0, 1]],
                                                                                                np.int32),
                        device=dev)
    nf = S.dos_n_freq(T, dt)
    ms = _timed(lambda: S.dos_spectra(vel, ind3, ntype, nf), reps=2)
    rec["dos"] = {"frames": T, "atoms": int(pos.shape[1]), "frequencies": nf, "ms": ms,
                  "alg_GBps": T * pos.shape[1] * 16 / (ms * 1e-3) / 1e9}
    del vel
    torch.cuda.empty_cache()
    return rec


def direct_dft_check(a, pos, counts, w, qsets, S_sel, geo, n_f, dt, nat):
    """S(Q,w) and S(Q) of the GPU forms against the FP64 oracle pipeline at
    the product set of seeded random Q: staging (mp_sqom56.f90:1115-1118),
    the direct DFT (dirft2d1's definition, :2530-2602, in 3D) of every frame,
    the species combine (:1338-1356) and the windowed time FFT with its
    normalisation (:1437-1527, :1576)."""
    from oracle import oracle as O

    t = time.perf_counter()
    T = pos.shape[0]
    K = int(np.prod([len(q) for q in qsets]))
    off = np.concatenate([[0], np.cumsum(counts)])
    F = np.zeros((T, 1, K), np.complex128)
    for f in range(T):
        pf = pos[f].cpu().numpy()
        parts = [O.stage(pf[off[j]:off[j + 1]], None, (0, 0, 0), np.eye(3), geo["tpe"], (0, 0, 0), 0, float(w[j]))
                 for j in range(len(counts))]
        x, y, z, c = [np.concatenate([p[k] for p in parts]) for k in range(4)]
        F[f, 0] = O.dirft_type1_prod(x, y, z, c, *qsets).ravel()
    Sd, Std = O.sqw_direct(F, n_int=T, win_kind=0, dt=dt, nat=nat, m_lo=0, n_f=n_f)
    out = {"q_points": K, "q_set": "product of %d x %d x %d seeded random k1, k2, k3" % tuple(len(q) for q in qsets),
           "frames": T, "eps": a.eps,
           "what": "rel l2 over K seeded Q x all m of S(Q,w), and over K Q of S(Q), vs the FP64 oracle pipeline "
                   "(staging + direct-DFT amplitudes of every frame + FP64 reduction)"}
    for k, (S, St) in S_sel.items():
        out["rel_l2_S_" + k] = float(np.linalg.norm(S - Sd) / np.linalg.norm(Sd))
        out["rel_l2_Stot_" + k] = float(np.linalg.norm(St - Std) / np.linalg.norm(Std))
    out["wall_s"] = time.perf_counter() - t
    return out


def c3_record(a, dev):
    """The precision sweep on one static frame of a 128^3 perovskite
    (10,485,760 atoms, 512 x 512 x 64 modes): FP32 and plain FP64 forms at
    eps 1e-3, 1e-4, 1e-6, 1e-9 (FP32 floors eps at 1e-7), time per transform
    and relative l2 error against the FP64 direct DFT at the 10,000 seeded
    modes of tests/golden/c3_dft_10k.npz (written by tests/golden/make_c3_dft.py
    from the same seeded frame)."""
    import torch

    from mp_tools_amd import synth
    from mp_tools_amd.sqw import SqwEngine

    g = np.load(os.path.join(ROOT, "tests", "golden", "c3_dft_10k.npz"))
    ncell = int(g["ncell"])
    pos, counts = synth.static_frame(ncell, sigma=0.05, seed=int(g["seed"]))
    p = pos[0, :, :3].astype(np.float64)
    same = bool(np.allclose([p.sum(), (p * p).sum(), p[::9973].sum()], g["checksums"], rtol=1e-12))
    N = tuple(int(v) for v in g["n_modes"])
    dpos = torch.tensor(pos, device=dev)
    idx = torch.tensor(_flat_index(N, g["k1s"], g["k2s"], g["k3s"]), device=dev)
    ref = g["F"].ravel()
    rec = {"workload": f"perovskite {ncell}^3 cells ({int(counts.sum())} atoms, 3 species, sigma 0.05 A), "
                       f"one frame, {N[0]}x{N[1]}x{N[2]} modes, neutron weights",
           "modes_checked": int(ref.size), "frame_matches_fixture": same, "eps": {}}
    for eps in (1e-3, 1e-4, 1e-6, 1e-9):
        row = {}
        for prec in ("fp32", "fp64"):
            eng = SqwEngine(3, N, counts, eps=eps, precision=prec, weights=g["weights"],
                            tpe=(2 * np.pi / ncell,) * 3)
            amp = eng.amplitudes(dpos)
            torch.cuda.synchronize()
            reps = 3
            t = time.perf_counter()
            for _ in range(reps):
                eng.amplitudes(dpos, out=amp)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) / reps * 1e3
            got = amp[0, idx].cpu().numpy()
            row[prec] = {"ms": ms, "rel_l2_vs_direct_dft": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
                         "kernel_width": eng.info()["w"], "eps_floor_warning": bool(eng.warning)}
            del amp, eng
            torch.cuda.empty_cache()
        row["fp32_over_fp64_speedup"] = row["fp64"]["ms"] / row["fp32"]["ms"]
        rec["eps"][f"{eps:g}"] = row
    return {"record": rec, "pos": (dpos, counts)}


def serve_record(a, dev, pos, counts, c3):
    """Serving (config 5) beside the step: 4 resident trajectories -- the
    step's 40^3 x 512 trajectory, the 128^3 static frame, a 40^3 x 64
    trajectory and a 120 A amorphous box x 64 frames -- and Poisson arrivals
    of map requests (species set, Q box 64^2x8 .. 256^2x32, 1-256 frames,
    neutron or seeded X-ray form factors) at three rates; throughput and
    p50/p99 latency per rate.  Two requests (neutron, X-ray) are served in
    the mix and kept for the oracle check of the CPU leg."""
    import torch

    from mp_tools_amd import synth
    from mp_tools_amd.serve import DONE, MapServer, poisson_load

    t0 = time.time()
    A = synth.A_LATTICE
    if c3 is None:
        p3, c3c = synth.static_frame(128, sigma=0.05, seed=20261205)
        p3 = torch.tensor(p3, device=dev)
    else:
        p3, c3c = c3
    p4, c4 = synth.phonon_trajectory(40, 64, dt_ps=0.01, n_phonons=64, rms=0.06, noise=0.01, seed=a.seed + 4,
                                     device=dev)
    p5, c5 = synth.amorphous_trajectory(120.0, 64, seed=a.seed + 5, device=dev)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    trajs = [dict(pos=pos, counts=counts, tpe=(2 * np.pi / 40,) * 3, qs=(2 * np.pi / (40 * A),) * 3,
                  b=0.1 * synth.B_NEUTRON_FM, nsp=3, T=pos.shape[0]),
             dict(pos=p3, counts=c3c, tpe=(2 * np.pi / 128,) * 3, qs=(2 * np.pi / (128 * A),) * 3,
                  b=0.1 * synth.B_NEUTRON_FM, nsp=3, T=1),
             dict(pos=p4, counts=c4, tpe=(2 * np.pi / 40,) * 3, qs=(2 * np.pi / (40 * A),) * 3,
                  b=0.1 * synth.B_NEUTRON_FM, nsp=3, T=64),
             dict(pos=p5, counts=c5, tpe=(2 * np.pi / 120.0,) * 3, qs=(2 * np.pi / 120.0,) * 3,
                  b=0.1 * synth.B_AL_U_FM, nsp=2, T=64)]
    srv = MapServer(hbm_budget=64 << 30, step_budget=8 << 30, max_active=64)
    for t in trajs:
        srv.add_trajectory(t["pos"], t["counts"], t["tpe"], t["qs"])
    poisson_load(srv, trajs, rate=50.0, duration_s=0.2, seed=a.seed)     # warm-up: engines, buffers
    rng = np.random.default_rng(a.seed + 7)
    xray = np.concatenate([rng.uniform(0.5, 10, (2, 4)), rng.uniform(0.5, 60, (2, 4)), rng.uniform(0, 1, (2, 1))],
                          axis=1)
    checked = [dict(traj=0, n_modes=(128, 128, 16), f0=100, T=48, mask=[1, 0, 1], n_f=24, weights=trajs[0]["b"]),
               dict(traj=3, n_modes=(64, 64, 32), f0=0, T=64, mask=[1, 1], n_f=32, xray=xray)]
    sweep, tks = [], []
    for k, rate in enumerate(a.serve_rates):
        if k < len(checked):
            tks.append(srv.submit(eps=1e-6, **checked[k]))
        nreq_target = a.serve_requests
        lat, qw, wall, nreq = poisson_load(srv, trajs, rate=rate, duration_s=nreq_target / rate,
                                           seed=a.seed + int(rate))
        sweep.append({"rate_req_s": rate, "requests": int(nreq), "throughput_req_s": len(lat) / wall,
                      "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                      "mean_queue_ms": float(np.mean(qw)), "wall_s": wall})
    srv.wait_idle()
    kept = []
    for r, tk in zip(checked, tks):
        st, _, lat_ms = srv.poll(tk)
        S, St = srv.result(tk)
        N = r["n_modes"]
        sets = _mode_set(np.random.default_rng(a.seed + 9), N, (6, 6, 4))
        idx = torch.tensor(_flat_index(N, *sets), device=dev)
        t = trajs[r["traj"]]
        kept.append(dict(req=r, sets=sets, done=(st == DONE), S=S[:, idx].cpu().numpy(), St=St[idx].cpu().numpy(),
                         pos=t["pos"][r["f0"]:r["f0"] + r["T"]].cpu().numpy(), counts=t["counts"], tpe=t["tpe"],
                         qs=t["qs"], nsp=t["nsp"]))
        srv.release(tk)
    stats = srv.stats()
    srv.close()
    best = max(sweep, key=lambda d: d["throughput_req_s"])
    rec = {"what": "Poisson arrivals of map requests against 4 resident trajectories (40^3x512, 128^3x1, 40^3x64 "
                   "perovskite; 120 A x64 amorphous); requests: species set, Q box 64^2x8..256^2x32, 1-256 frames, "
                   "neutron or X-ray weights, eps 1e-6; HBM-budget admission, continuous batching",
           "sweep": sweep, "best_throughput_req_s": best["throughput_req_s"],
           "p50_ms_at_lowest_rate": sweep[0]["p50_ms"], "p99_ms_at_lowest_rate": sweep[0]["p99_ms"],
           "server": stats, "generation_s": gen_s}
    return {"record": rec, "checked": kept}


def served_map_check(srv):
    """The two requests served in the mix against the FP64 oracle pipeline
    at a product set of 144 seeded Q: per-species staging + direct DFT, the
    request's weights (neutron b, or X-ray f_s(|Q|) = sum_k a_k exp(-b_k
    |Q|^2 / 16 pi^2) + c evaluated in FP64 and rounded to FP32 as the server
    stores it), and the windowed time FFT."""
    from oracle import oracle as O

    out = []
    for k in srv["checked"]:
        r, sets = k["req"], k["sets"]
        act = [j for j in range(k["nsp"]) if r["mask"][j]]
        off = np.concatenate([[0], np.cumsum(k["counts"])])
        T = k["pos"].shape[0]
        F = np.zeros((T, len(act), int(np.prod([len(q) for q in sets]))), np.complex128)
        for t in range(T):
            for a_, j in enumerate(act):
                x, y, z, c = O.stage(k["pos"][t, off[j]:off[j + 1]], None, (0, 0, 0), np.eye(3), k["tpe"],
                                     (0, 0, 0), 0, 1.0)
                F[t, a_] = O.dirft_type1_prod(x, y, z, c, *sets).ravel()
        nat = float(np.sum(k["counts"]))
        if "xray" in r:
            K3, K2, K1 = np.meshgrid(*[s * q for s, q in zip(sets[::-1], k["qs"][::-1])], indexing="ij")
            q2 = (K1 ** 2 + K2 ** 2 + K3 ** 2).ravel()
            x = r["xray"]
            wtab = np.stack([(x[s, 8] + sum(x[s, i] * np.exp(-x[s, 4 + i] * q2 / (16 * np.pi ** 2))
                                            for i in range(4))).astype(np.float32).astype(np.float64) for s in act])
            Sr, Str = O.sqw_direct(F, wtab=wtab, n_int=T, dt=0.01, nat=nat, n_f=r["n_f"])
        else:
            Sr, Str = O.sqw_direct(F, wsc=np.asarray(r["weights"])[act], n_int=T, dt=0.01, nat=nat, n_f=r["n_f"])
        out.append({"weights": "xray" if "xray" in r else "neutron", "modes": list(r["n_modes"]), "frames": T,
                    "done": k["done"], "q_points": int(F.shape[2]),
                    "rel_l2_S": float(np.linalg.norm(k["S"] - Sr) / np.linalg.norm(Sr)),
                    "rel_l2_Stot": float(np.linalg.norm(k["St"] - Str) / np.linalg.norm(Str))})
    return out


def batch_bench(a):
    """64 uneven jobs (half perovskite supercells of side 8-48 cells with
    phonon trajectories, half amorphous Al/U-like boxes of side 30-120 A with a
    2.2 A hard core; 1-64 frames; 32^3 .. 256x256x32 modes; FP32, eps 1e-6):
    spread tiles autotuned per geometry, then all jobs in one launch per phase
    (S(Q,w) reductions included), captured in a hipGraph and replayed."""
    import torch

    from mp_tools_amd import synth
    from mp_tools_amd.sqw import SqwBatch, SqwEngine, autotune_tiles

    dev = torch.device("cuda:0")
    t0 = time.time()
    jobs = synth.batch_jobs(n_jobs=a.jobs, seed=a.seed + 2, device=dev, eps=1e-6)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    # autotune every distinct geometry once (cached for the batch's engines)
    t0 = time.time()
    tuned = {}
    for kw, pos, red in jobs:
        key = tuple(kw["n_modes"])
        if key in tuned:
            continue
        k2 = {k: v for k, v in kw.items() if k not in ("dim", "n_modes", "species_counts")}
        tuned[key] = autotune_tiles(pos, kw["dim"], kw["n_modes"], kw["species_counts"], **k2)
    tune_s = time.time() - t0
    b = SqwBatch([kw for kw, _, _ in jobs])
    pos = [p for _, p, _ in jobs]
    reds = [r for _, _, r in jobs]
    frames = sum(p.shape[0] for p in pos)
    pts = sum(p.shape[0] * p.shape[1] for p in pos)
    b.set_timing(True)
    out = b.run(pos, reds, use_graph=True)      # eager (allocates), stage timings
    torch.cuda.synchronize()
    stages = {k: {"ms": v[0], "launches": v[1]} for k, v in b.timing().items() if v[1]}
    b.run(pos, reds, out=out, use_graph=True)   # capture + first replay
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        b.run(pos, reds, out=out, use_graph=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        b.run(pos, reds, out=out, use_graph=True)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / a.steps
    # the same jobs one engine after another (what a caller without batching
    # does): per job a warm-up call, then the timed call
    sec_seq, same = 0.0, True
    for j, (kw, p_, r_) in enumerate(jobs):
        kw = dict(kw)
        e = SqwEngine(kw.pop("dim"), kw.pop("n_modes"), kw.pop("species_counts"), **kw)
        a1 = torch.empty_like(out[0][j])
        e.amplitudes(p_, out=a1)
        e.reduce(a1, **r_)
        torch.cuda.synchronize()
        t = time.perf_counter()
        e.amplitudes(p_, out=a1)
        e.reduce(a1, **r_)
        torch.cuda.synchronize()
        sec_seq += time.perf_counter() - t
        same = same and bool(torch.equal(a1, out[0][j]))
        del e, a1
    torch.cuda.empty_cache()
    res = {
        "metric": "batched S(Q,w) jobs: frames/s over all jobs (FP32, eps 1e-6, one launch per phase, hipGraph)",
        "value": frames / sec, "unit": "frames/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded perovskite phonon trajectories and RSA amorphous boxes)",
        "config": {"workload": f"{a.jobs} uneven jobs: perovskite side 8-48 cells / amorphous side 30-120 A, "
                               f"1-64 frames, modes 32^3..256x256x32", "jobs": a.jobs, "frames": frames,
                   "points": pts, "seed": a.seed + 2},
        "jobs_per_s": a.jobs / sec, "nu_pts_per_s": pts / sec,
        "one_by_one_ms": sec_seq * 1e3, "batch_speedup_vs_one_by_one": sec_seq / sec,
        "batch_equals_one_by_one": bool(same),
        "stages_eager": stages, "generation_s": gen_s, "autotune_s": tune_s,
        "tuned_tiles": {"x".join(map(str, k)): [v[0], v[1], v[2]] for k, v in tuned.items()},
    }
    print(json.dumps(res))


def serve_bench(a):
    """Serving loop: 4 device-resident trajectories (the S(Q,w) benchmark's
    40^3 perovskite, 512 frames; a static 128^3 perovskite frame; a 40^3
    perovskite job of 64 frames and a 120 A amorphous box of 64 frames from
    the batched workload); Poisson arrivals of map requests (species set,
    Q box 64^2x8 .. 256^2x32, 1-256 frames, neutron or seeded X-ray form
    factors) at each rate of the sweep; HBM-budget admission and continuous
    batching in the native server; throughput and p50/p99 latency."""
    import torch

    from mp_tools_amd import synth
    from mp_tools_amd.serve import MapServer, poisson_load

    dev = torch.device("cuda:0")
    t0 = time.time()
    p2, c2 = synth.phonon_trajectory(40, 512, dt_ps=0.01, n_phonons=64, rms=0.06, noise=0.01, seed=a.seed,
                                     device=dev)
    p3, c3 = synth.static_frame(128, sigma=0.05, seed=a.seed + 3)
    p3 = torch.tensor(p3, device=dev)
    p4, c4 = synth.phonon_trajectory(40, 64, dt_ps=0.01, n_phonons=64, rms=0.06, noise=0.01, seed=a.seed + 4,
                                     device=dev)
    p5, c5 = synth.amorphous_trajectory(120.0, 64, seed=a.seed + 5, device=dev)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    A = synth.A_LATTICE
    trajs = [dict(pos=p2, counts=c2, tpe=(2 * np.pi / 40,) * 3, qs=(2 * np.pi / (40 * A),) * 3,
                  b=0.1 * synth.B_NEUTRON_FM, nsp=3, T=512),
             dict(pos=p3, counts=c3, tpe=(2 * np.pi / 128,) * 3, qs=(2 * np.pi / (128 * A),) * 3,
                  b=0.1 * synth.B_NEUTRON_FM, nsp=3, T=1),
             dict(pos=p4, counts=c4, tpe=(2 * np.pi / 40,) * 3, qs=(2 * np.pi / (40 * A),) * 3,
                  b=0.1 * synth.B_NEUTRON_FM, nsp=3, T=64),
             dict(pos=p5, counts=c5, tpe=(2 * np.pi / 120.0,) * 3, qs=(2 * np.pi / 120.0,) * 3,
                  b=0.1 * synth.B_AL_U_FM, nsp=2, T=64)]
    srv = MapServer(hbm_budget=64 << 30, step_budget=8 << 30, max_active=64)
    for t in trajs:
        srv.add_trajectory(t["pos"], t["counts"], t["tpe"], t["qs"])
    # warm-up: one request per trajectory (engines, buffers)
    poisson_load(srv, trajs, rate=50.0, duration_s=0.1, seed=a.seed)
    sweep = []
    for rate in (5.0, 20.0, 50.0, 100.0, 200.0):
        nreq_target = 40
        lat, qw, wall, nreq = poisson_load(srv, trajs, rate=rate, duration_s=nreq_target / rate,
                                           seed=a.seed + int(rate))
        sweep.append({"rate_req_s": rate, "requests": int(nreq), "throughput_req_s": len(lat) / wall,
                      "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                      "mean_queue_ms": float(np.mean(qw)), "wall_s": wall})
        print(json.dumps(sweep[-1]), file=sys.stderr, flush=True)
    best = max(sweep, key=lambda d: d["throughput_req_s"])
    res = {
        "metric": "served S(Q,w) maps: throughput (req/s) and p50/p99 latency under Poisson load",
        "value": best["throughput_req_s"], "unit": "req/s", "n_gpus": 1, "steps": len(sweep), "warmup": 1,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded trajectories; seeded request stream)",
        "config": {"workload": "4 resident trajectories (40^3x512, 128^3x1, 40^3x64 perovskite; 120 A x64 "
                               "amorphous); requests: species set, Q box 64^2x8..256^2x32, 1-256 frames, "
                               "neutron or X-ray weights, eps 1e-6", "requests_per_rate": 40},
        "sweep": sweep, "server": srv.stats(), "generation_s": gen_s,
    }
    srv.close()
    print(json.dumps(res))


def copy_calibration(mib):
    """Five launches of an elementwise kernel that reads mib MiB and writes
    mib MiB (torch.neg on float32): the known-size kernel the HBM counters are
    scaled by."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t) / 5
    print(json.dumps({"copy_mib": mib, "bytes_per_launch": 2 * n * 4, "ms_per_launch": el * 1e3,
                      "GBps": 2 * n * 4 / el / 1e9}))


def sql_bench(a):
    """mp_sql on the reference's own Al92U08 snapshot (30,000 atoms, 79.8 A
    box): Q range 20 1/A = 509^3 modes per species, X-ray weights.  A step is
    one snapshot: 2 species' 3D NUFFTs (eps 1e-6) + the |Q| histogram; the
    output functions (host, 2001 bins) are outside the timed region.  The
    value is the FP32 form (the reference's amplitudes are REAL(4) complex;
    FP64 histogram sums), with the FP64 form and the Monte Carlo integration
    (16 x 10^6 samples per snapshot, FP32) beside it.  The CPU leg times the
    host restatement of the same steps (ES spreading in C, scipy FFT, numpy
    binning) at Q range 5 1/A (129^3 modes) as its sample."""
    import torch

    from mp_tools_amd import mpdat, sql
    from oracle import oracle as O
    snap = mpdat.read_dat(os.path.join(ROOT, "tests", "golden", "al_u", "Al_U_large_n0001.dat"))
    counts = snap.species_counts
    g = sql.geometry(snap.a_par, snap.a_cell_inv, snap.header["N_ROW"], snap.n_tot, 20.0, 0.01)
    torch.cuda.set_device(0)
    frames = [snap.at_pos] * a.steps

    def timed(precision, **kw):
        for _ in range(max(1, a.warmup)):
            sql.histogram([snap.at_pos], counts, g, precision=precision, **kw)
        tm = {}
        t0 = time.time()
        sql.histogram(frames, counts, g, precision=precision, timing=tm, **kw)
        return tm["ms"] / a.steps, time.time() - t0, tm

    ms, wall, _ = timed("fp32")
    ms64, _, _ = timed("fp64")
    ms_mc, _, tmc = timed("fp32", integration="mc", n_h=16, seed=1)
    # host restatement on a bounded sample: Q range 5
    gs = sql.geometry(snap.a_par, snap.a_cell_inv, snap.header["N_ROW"], snap.n_tot, 5.0, 0.01)
    cores = os.cpu_count() or 1
    t1 = time.time()
    off = np.concatenate([[0], np.cumsum(counts)])
    amps = []
    for s_ in range(len(counts)):
        r = snap.at_pos[off[s_]:off[s_ + 1], :3].astype(np.float32)
        x, y, z = [(r[:, d] * gs.dq[d]).astype(np.float64) for d in range(3)]
        amps.append(O.cpu_nufft_type1(x, y, z, np.ones(len(r), np.complex128), gs.n_qq, eps=1e-6, iflag=-1,
                                      workers=cores).ravel())
    k = [np.arange(n) - n // 2 for n in gs.n_qq[::-1]]
    kz, ky, kx = np.meshgrid(*k, indexing="ij")
    qn = np.sqrt((kx * gs.dq[0]) ** 2 + (ky * gs.dq[1]) ** 2 + (kz * gs.dq[2]) ** 2).ravel()
    sel = qn <= gs.q_range
    b = np.rint(qn[sel] / gs.q_step).astype(np.int64)
    for i_ in range(len(amps)):
        for j_ in range(i_ + 1):
            np.bincount(b, (amps[i_][sel] * np.conj(amps[j_][sel])).real, minlength=gs.n_hist)
    cpu_s = time.time() - t1
    res = {
        "metric": "mp_sql snapshots/s (Al92U08, 30,000 atoms, Q range 20/A = 509^3 modes x 2 species, FP32 NUFFT + |Q| histogram)",
        "value": 1000.0 / ms, "unit": "snapshots/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "fp64_form": {"value": 1000.0 / ms64, "ms_per_step": ms64},
        "monte_carlo": {"value": 1000.0 / ms_mc, "ms_per_step": ms_mc, "samples_per_snapshot": 16 * sql.N_MC,
                        "skips_last_snapshot": list(tmc["skips"]),
                        "note": "Gauss sum replaced by 16 chains of 10^6 Philox-drawn modes (mp_sql j_acc = 2)"},
        "data": "the reference's example snapshot examples/amorphous_Al92U08/data/Al_U_large_n0001.dat (tests/golden/al_u)",
        "config": {"workload": "mp_sql Al/U Q_range 20, q_step 0.01, X-ray", "n_qq": list(g.n_qq), "n_hist": g.n_hist},
        "wall_s": wall,
        "reference_published": {"value_s_per_snapshot": 11.5, "machine": "Apple M1, 4 OpenMP threads",
                                "source": "examples/amorphous_Al92U08/aaa_readme_amorphous_Al92U08.txt (Q_range 20, AVERAGE column)"},
        "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "snapshots/s at the sample size", "cores": cores, "kind": "port",
                         "sample": f"one snapshot at Q range 5 ({gs.n_qq[0]}^3 modes): 2 species' NUFFTs, host ES spreading + scipy FFT"},
    }
    print(json.dumps(res))


def cpu_leg(a, pos, counts, w, N, amp_gpu, geo):
    """Host FINUFFT-method restatement (the oracle's FP64 ES spreading in
    C/OpenMP with kernel values once per point, scipy.fft, cached
    deconvolution factors) of the first F frames: one transform per frame
    over all atoms with strengths w_s c_j (the same merge the GPU uses),
    then the time reduction of those frames; frames/s = F / wall, on all host
    cores and on one core."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    off = np.concatenate([[0], np.cumsum(counts)])
    import scipy.fft as sfft

    def frames_per_s(F, threads):
        O.set_threads(threads)
        pn = pos[:F].cpu().numpy()
        t = time.perf_counter()
        C = np.zeros((F, int(np.prod(N))), np.complex128)
        for f in range(F):
            parts = [O.stage(pn[f, off[j]:off[j + 1]], None, (0, 0, 0), np.eye(3), geo["tpe"], (0, 0, 0), 0,
                             float(w[j])) for j in range(len(counts))]
            x, y, z, c = [np.concatenate([p[k] for p in parts]) for k in range(4)]
            C[f] = O.cpu_nufft_type1(x, y, z, c, N, eps=a.eps, workers=threads).ravel()
        t_nufft = time.perf_counter() - t
        t = time.perf_counter()
        win = 0.5 * (1 - np.cos(2 * np.pi * np.arange(F) / max(F - 1, 1)))
        Cp = (C - C.mean(0)) * win[:, None]
        S = np.abs(sfft.fft(Cp, axis=0, workers=threads)) ** 2
        Sm = S.sum(axis=0)
        t_red = time.perf_counter() - t
        del S, Sm
        return F / (t_nufft + t_red), t_nufft, t_red, C

    F = max(1, a.cpu_frames)
    fps, t_nufft, t_red, C = frames_per_s(F, cores)
    g = amp_gpu[:F].reshape(F, -1)
    err = float(np.linalg.norm(g - C) / np.linalg.norm(C))
    del C
    F1 = max(1, a.cpu_frames_1core)
    fps1, t1n, t1r, _ = frames_per_s(F1, 1)
    O.set_threads(cores)
    return {"value": fps, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"first {F} of the {a.frames} frames: host FINUFFT-method restatement (FP64 ES "
                      f"spreading in C/OpenMP, kernel values once per point; scipy.fft; cached "
                      f"deconvolution), one transform per frame over all 3 species with strengths "
                      f"w_s c_j, then the windowed time FFT of those frames",
            "precision": "fp64", "wall_s": t_nufft + t_red, "nufft_s": t_nufft, "reduce_s": t_red,
            "one_core": {"value": fps1, "unit": "frames/s", "cores": 1, "frames": F1, "wall_s": t1n + t1r},
            "rel_l2_gpu_fp32_amplitudes_vs_cpu": err}


if __name__ == "__main__":
    main()
