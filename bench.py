"""Benchmark: SC+PC-NEVPT2 evaluations per second on one MI355X.

One step = one full NEVPT2 evaluation (or, with --concurrency S, a batch of S concurrent
evaluations through the request loop) on device-resident inputs: partial traces
of the packed 4-RDM, every Koopmans matrix (including the O(n^9) A-matrices),
canonical orbitals, Cholesky integral assembly and the eight perturber classes
in SC and PC form.  Inputs: a seeded synthetic CI vector (truncated to its 2e5
largest determinants for CAS(14,14)) from which the RDMs are built on the GPU,
and seeded Cholesky ERIs; both are resident in HBM before the timed region.

The default workload is CAS(14,14), nc=30, nv=400.  Three precision forms are
timed: the plain FP64 path and the mixed forms (4-RDM contractions in FP32, and
in split-BF16x3 on the bf16 matrix cores with FP32 accumulation).  `value` is
the throughput of the fastest form whose per-class energies stay within 1 uEh
of FP64 on this workload; the other forms are reported beside it with their
per-class |dE| against FP64 and the class norms that amplify it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--prec all|fp64|fp32|bf16x3]
                    [--concurrency S]
    python bench.py --config c5 [--rates 5,10,20] [--slots 4]     (serving: Poisson arrivals, p50/p99)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {   # name: (nact, ncore, nvirt, roots, determinants kept, description)
    "c1": (6, 10, 40, 1, None, "CAS(6,6) nc=10 nv=40 single state"),
    "c2": (10, 20, 200, 1, None, "CAS(10,10) nc=20 nv=200 single state"),
    "c3": (14, 30, 400, 1, 200000, "CAS(14,14) nc=30 nv=400 single state, CI truncated to its 2e5 largest determinants"),
    "c4": (12, 25, 300, 6, None, "CAS(12,12) nc=25 nv=300, 6 roots (QD-NEVPT2)"),
}
# CPU oracle sample per workload: bamat (a,c) pairs run (None: one per host thread, so that the
# threaded pair loop runs one full wave), fraction of each class's outer external loop
CPU_SAMPLE = {"c1": (5, 0.05), "c2": (5, 0.05), "c3": (None, 0.01), "c4": (None, 0.01)}
METRIC = "SC+PC-NEVPT2 evals/sec CAS(12-14) mixed vs FP64 at |ΔE|≤1µEh; p50/p99 lat"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 157.3e12}   # bf16x3 stands in for FP32 arithmetic
HBM_PEAK = 8.0e12
TOL_EH = 1.0e-6
CLASS_NAMES = ["(0)", "(+1)", "(-1)", "(+2)", "(-2)", "(+1)'", "(-1)'", "(0)'"]
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC = {   # (kernel, config, precision) -> committed rocprofv3 counter summary holding its HBM bytes
    ("contractions", "c3", "fp64"): ("e4_contractions_cas14_fp64_pmc.md", "traffic_bytes_per_eval:"),
    ("e4_build", "c3", "fp64"): ("e4_contractions_cas14_fp64_pmc.md", "e4_build_traffic_bytes_per_eval:"),
    ("trd_col", "c3", "fp64"): ("trd_col_cas14_fp64_pmc.md", "traffic_bytes_per_eval:"),
}


def contraction_flops(n, M):
    """Algorithmic flops of bamat's four 4-RDM contractions: 4 x 2 n^9 per state."""
    return 8.0 * n ** 9 * M


def contraction_bytes(n, M, esize):
    """Algorithmic HBM bytes of the two contraction passes per evaluation: E4 (n^8 elements of
    esize bytes) read once by each pass, and the four n^6 FP64 outputs written."""
    return M * (2.0 * n ** 8 * esize + 4.0 * 8 * n ** 6)


def trd_bytes(N, nb):
    """Algorithmic bytes of one tridiagonal reduction (k_trd_col launches k = 0..N-2): launch k reads
    and writes the trailing block of order m = N-k-1 of each of the nb matrices once."""
    m = np.arange(1, N, dtype=np.float64)
    return 16.0 * nb * float((m * m).sum())


def band_flops(N, nb):
    """Algorithmic flops of the reduction to band form (half-bandwidth 64): per panel of m trailing rows,
    W = A22 Y and the rank-128 update A22 -= Y Z^T + Z Y^T, 6 * 64 m^2, plus the CholeskyQR2 and
    reflector products on the m x 64 panel, about 10 * 64^2 m."""
    m = np.arange(N - 64, 64, -64, dtype=np.float64)
    return nb * float((6.0 * 64 * m * m + 10.0 * 64 * 64 * m).sum())


def e4_build_bytes(n, M, esize):
    """E4 (n^8 elements) written once, the packed Gamma4 (nwords * n^4) and the 3-RDM (n^6, read and
    transposed once) read once, per state."""
    nw = n * (n + 1) * (n + 2) * (n + 3) // 24
    return M * (esize * n ** 8 + 8 * nw * n ** 4 + 3 * 8 * n ** 6)


def measured_traffic(kernel, config, prec):
    """HBM bytes per evaluation of a kernel family from the committed rocprofv3 counter runs
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), for the workload they were taken on."""
    ent = TRAFFIC.get((kernel, config, prec))
    if not ent:
        return None
    path = os.path.join(PROFILES, ent[0])
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(ent[1]):
            return float(line.split(":")[1])
    return None


def cpu_baseline(P, pairs=5, frac=0.05):
    """Oracle timed on the host cores on a bounded sample of the same workload.

    Stage 1: every Koopmans build in full except bamat, whose (a,c) outer loop is
    run on `pairs` of its n^2 pairs and scaled; stage 2: every class with its
    outer external loop sampled at `frac` and scaled (oracle.stage2 ``frac``).
    """
    from oracle import stage1 as S1, stage2 as S2
    n, nc, M = P.nact, P.ncore, P.metat
    t0 = time.perf_counter()
    aj, atwo = S1.get_aj_atwo(nc, n, P.norb, P.B, P.oneint)
    koop = {"aj": aj, "atwo": atwo}
    koop["1-IPKO"], koop["1-EAKO"] = S1.koopE(n, nc, M, atwo, P.taa, P.dal, aj)
    koop["2-HRDM"] = S1.hrdm(n, M, P.taa, P.dal)
    koop["2-IPKO"], koop["2-EAKO"] = S1.koop2E(n, nc, M, atwo, P.daaa, P.taa, P.dal, aj)
    koop["gk0pa"], koop["gk0pd"], koop["gk0pf"] = S1.koopman0pE(n, nc, M, atwo, P.daaa, P.taa, P.dal, aj)
    koop["bmatIP"], koop["bmatEA"] = S1.bbmat(n, nc, M, atwo, P.daaa, P.taa, P.dal, aj)
    koop["cmatIP"], koop["cmatEA"] = S1.bcmat(n, nc, M, atwo, P.daaa, P.taa, P.dal, aj)
    koop["dmatIP"], koop["dmatEA"] = S1.bdmat(n, nc, M, atwo, P.taa, P.dal, aj)
    t_small = time.perf_counter() - t0
    npairs = n * n
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))
    k = max(1, min(pairs if pairs else threads, npairs))
    t0 = time.perf_counter()
    am, at = S1.bamat(n, nc, M, atwo, P.ro4, P.daaa, P.taa, P.dal, aj, pair_limit=k)
    t_bamat = (time.perf_counter() - t0) * npairs / k
    koop["amatIP"], koop["amatEA"] = am, at        # partial; stage-2 timing does not depend on values
    _, _, t_s2 = S2.state_energies_timed(P, koop, 0, frac=frac)
    t_total = t_small + t_bamat + t_s2
    sample = (f"stage 1: full except bamat on {k}/{npairs} (a,c) pairs scaled x{npairs / k:.0f}; stage 2: each "
              f"class's outer external loop on a {frac:.0%} sample, scaled (the metric and K' eigen-solves in full)")
    return {"value": 1.0 / t_total, "unit": "evals/s", "cores": threads,
            "kind": "port", "sample": sample,
            "seconds_per_eval": {"stage1_small": t_small, "bamat": t_bamat, "stage2": t_s2, "total": t_total}}


def host_rdms(P, ev):
    """The oracle's inputs: ro4 and its partial traces as built on the GPU, copied to the host."""
    n, M = P.nact, P.metat
    nw = n * (n + 1) * (n + 2) * (n + 3) // 24
    P.ro4 = ev.get("ro4", (nw, n, n, n, n, M))
    P.daaa = ev.get("3-RDM", (M,) + (n,) * 6)
    P.taa = ev.get("2-RDM", (M,) + (n,) * 4)
    P.dal = ev.get("1-RDM", (M, n, n))


def time_form(evs, prec, steps, warmup, S):
    """Unloaded latency, per-phase device times and the timed steps of one precision form."""
    from nevpt2_amd.evaluator import PHASES
    from nevpt2_amd.serve import Server
    for ev in evs:
        for _ in range(warmup):
            ev.run(prec)
    lat1, phases, kred, kform = [], np.zeros(len(PHASES)), [], 0
    for _ in range(steps):
        t = time.perf_counter()
        res = evs[0].run(prec)
        lat1.append((time.perf_counter() - t) * 1e3)
        phases += np.array(list(evs[0].timings().values()))
        pt = evs[0].pc_timings()
        kred.append(pt[:, 0].sum())
        kform = int(pt[:, 3].max())
    phases /= steps
    latS = []
    if S == 1:         # the timed steps: one evaluation each, on the calling thread
        t0 = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            evs[0].run(prec)
            latS.append((time.perf_counter() - t) * 1e3)
        elapsed = time.perf_counter() - t0
    else:              # each step submits S evaluations to the request loop and waits for all of them
        with Server(slots=S) as srv:
            for ev in evs:                                 # one untimed round through the server
                srv.wait(srv.submit_ctx(ev, prec))
            t0 = time.perf_counter()
            for _ in range(steps):
                ids = [srv.submit_ctx(ev, prec) for ev in evs]
                for rid in ids:
                    _, _, lat, _ = srv.wait(rid)
                    latS.append(lat)
            elapsed = time.perf_counter() - t0
    return {"res": res, "elapsed": elapsed, "value": S * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
            "phases": dict(zip(PHASES, [float(x) for x in phases])), "kred_ms": float(np.mean(kred)), "kred_form": kform,
            "latency_ms": {"p50": float(np.percentile(latS, 50)), "p99": float(np.percentile(latS, 99)),
                           "unloaded_p50": float(np.percentile(lat1, 50)),
                           "unloaded_p99": float(np.percentile(lat1, 99))}}


SERVE_METRIC = "SC+PC-NEVPT2 evals/sec under Poisson arrivals (request loop); p50/p99 latency"


def _sub_problem(base, nc, nv, ncore_max, n):
    """Orbital sub-block of a seeded problem with ncore_max core and many virtual orbitals: the first nc core
    orbitals, the active ones and the first nv virtual ones, with the first 3 norb Cholesky vectors."""
    idx = np.r_[np.arange(nc), ncore_max + np.arange(n), ncore_max + n + np.arange(nv)]
    norb = len(idx)
    hi = np.maximum(idx[:, None], idx[None, :])
    lo = np.minimum(idx[:, None], idx[None, :])
    tri_old = hi * (hi + 1) // 2 + lo
    r, c = np.tril_indices(norb)
    cols = tri_old[r, c]                        # new pair (r >= c) in the reference's packed order
    order = np.argsort(r * (r + 1) // 2 + c)
    cols = cols[order]
    B = np.asfortranarray(base["B"][: 3 * norb, cols])
    return B, np.ascontiguousarray(base["oneint"][cols])


def run_serving(args):
    """256 requests, Poisson arrivals at each swept rate, through the request loop (nv2_server_*): each request
    is a host-data NEVPT2 evaluation (inputs uploaded on admission into a pooled or new context)."""
    from types import SimpleNamespace
    from nevpt2_amd import synth
    from nevpt2_amd.evaluator import Evaluator
    from nevpt2_amd.serve import Server
    rng = np.random.default_rng(2024)
    NREQ, NDISTINCT, NCMAX, NVMAX = 256, 32, 30, 300
    print("[bench c5] building the RDMs of each active size on the GPU (3 roots) ...", file=sys.stderr, flush=True)
    base = {}
    for n in (6, 8, 10, 12):
        Pn = synth.Problem(n, NCMAX, NVMAX, seed=n, nroots=3, host_rdms=False)
        ev = Evaluator.for_problem(Pn, rdms="gpu")
        nw = n * (n + 1) * (n + 2) * (n + 3) // 24
        base[n] = {"C": Pn.C, "B": Pn.B, "oneint": Pn.oneint, "ro4": ev.get("ro4", (nw, n, n, n, n, 3)),
                   "t3": ev.get("3-TRDM", (3,) + (n,) * 6), "t2": ev.get("2-TRDM", (3,) + (n,) * 4),
                   "t1": ev.get("1-TRDM", (3, n, n))}
        ev.close()
    # the distinct problems, drawn from the workload's distribution
    probs = []
    for _ in range(NDISTINCT):
        n = int(rng.choice([6, 8, 10, 12]))
        M = int(rng.integers(1, 4))
        nc = int(rng.integers(5, NCMAX + 1))
        nv = int(rng.integers(50, NVMAX + 1))
        b = base[n]
        B, one = _sub_problem(b, nc, nv, NCMAX, n)
        npr = M * (M - 1) // 2
        P = SimpleNamespace(ncore=nc, nact=n, nvirt=nv, metat=M, nele=n, B=B, oneint=one,
                            ro4=np.asfortranarray(b["ro4"][..., :M]),
                            daaaoff=np.asfortranarray(b["t3"][:npr]) if M > 1 else None,
                            taaoff=np.asfortranarray(b["t2"][:npr]) if M > 1 else None,
                            daloff=np.asfortranarray(b["t1"][:npr]) if M > 1 else None)
        probs.append(P)
    picks = rng.integers(0, NDISTINCT, NREQ)
    rates = [float(r) for r in args.rates.split(",")]
    results = []
    with Server(slots=args.slots) as srv:
        for P in probs:                        # warm the context pool: one request of each distinct problem
            srv.wait(srv.submit(P))
        for rate in rates:
            gaps = rng.exponential(1.0 / rate, NREQ)
            print(f"[bench c5] {NREQ} requests at {rate:g}/s ...", file=sys.stderr, flush=True)
            ids = []
            t0 = time.perf_counter()
            t_next = t0
            for k in range(NREQ):
                t_next += gaps[k]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids.append(srv.submit(probs[picks[k]]))
            lat = []
            for rid in ids:
                _, _, l, _ = srv.wait(rid)
                lat.append(l)
            elapsed = time.perf_counter() - t0
            results.append({"rate_per_s": rate, "evals_per_s": NREQ / elapsed,
                            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                           "mean": float(np.mean(lat))}})
            print(f"[bench c5]   {NREQ / elapsed:.2f} evals/s, p50 {np.percentile(lat, 50):.1f} ms, "
                  f"p99 {np.percentile(lat, 99):.1f} ms", file=sys.stderr, flush=True)
    best = max(results, key=lambda r: r["evals_per_s"])
    shapes = sorted({(P.nact, P.metat) for P in probs})
    out = {"metric": SERVE_METRIC, "value": best["evals_per_s"], "unit": "evals/s", "n_gpus": 1,
           "steps": NREQ, "warmup": NDISTINCT, "ms_per_step": 1e3 / best["evals_per_s"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: per active size one seeded CI vector set (3 roots; RDMs built on the GPU) and one seeded "
                   "integral generator with 30 core / 300 virtual orbitals, of which each request takes its first nc core "
                   "and nv virtual orbitals and 3 norb Cholesky vectors; 32 distinct requests drawn from the workload's "
                   "distribution, 256 arrivals picked among them",
           "config": {"workload": "serving: 256 requests, Poisson arrivals, n in {6,8,10,12}, 1-3 roots (QD couplings), "
                                  "nc in [5,30], nv in [50,300], host-data requests uploaded on admission",
                      "slots": args.slots, "precision": "fp64", "distinct_shapes_nact_roots": shapes},
           "rates": results,
           "note": "FP64: the mixed forms do not hold 1 uEh on these intruder-laden random CI vectors"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS) + ["c5"])
    ap.add_argument("--rates", default="5,10,20", help="c5: Poisson arrival rates (requests per second)")
    ap.add_argument("--slots", type=int, default=4, help="c5: request-loop slots")
    ap.add_argument("--prec", default="all", choices=["all", "both", "fp64", "fp32", "bf16x3"])
    ap.add_argument("--concurrency", type=int, default=1,
                    help="evaluations in flight per step through the request loop (one resident context each); "
                         "1 runs them one after another on the calling thread")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path would shard by external "
              f"orbital index, each GPU holding the RDMs and its slice of the external orbitals", file=sys.stderr)
        sys.exit(2)
    if args.config == "c5":
        run_serving(args)
        return
    from nevpt2_amd import synth
    from nevpt2_amd.evaluator import Evaluator
    n, nc, nv, M, keep, desc = CONFIGS[args.config]
    S = max(1, args.concurrency)
    print(f"[bench] generating {desc} ...", file=sys.stderr, flush=True)
    P = synth.Problem(n, nc, nv, seed=1, nroots=M, keep=keep, host_rdms=False)
    t = time.perf_counter()
    evs = [Evaluator.for_problem(P, rdms="gpu") for _ in range(S)]   # each request: its own context and inputs in HBM
    t_setup = (time.perf_counter() - t) / S
    print(f"[bench] RDMs built on the GPU from the CI vector(s), integrals uploaded: {t_setup:.2f} s per context",
          file=sys.stderr, flush=True)
    forms = {}
    for name, prec in (("fp64", 0), ("fp32", 1), ("bf16x3", 2)):
        if args.prec in ("all", name) or (args.prec == "both" and name != "bf16x3"):
            print(f"[bench] timing the {name} form ...", file=sys.stderr, flush=True)
            forms[name] = time_form(evs, prec, args.steps, args.warmup, S)
    ref = forms["fp64"]["res"] if "fp64" in forms else evs[0].run(0)
    for name in ("fp32", "bf16x3"):          # the mixed forms against the FP64 path on the same resident inputs
        if name not in forms:
            continue
        got = forms[name]["res"]
        d = np.abs(got - ref)
        t64 = forms["fp64"]["phases"]["E4 contractions"] if "fp64" in forms else None
        tmx = forms[name]["phases"]["E4 contractions"]
        forms[name]["vs_fp64"] = {
            "max_abs_dE_sc": float(d[:, :, 1].max()), "max_abs_dE_pc": float(d[:, :, 3].max()),
            "abs_dE_total_sc": float(abs(got[:, :, 1].sum() - ref[:, :, 1].sum())),
            "abs_dE_total_pc": float(abs(got[:, :, 3].sum() - ref[:, :, 3].sum())), "unit": "Eh",
            # per class (reference order): |dE| SC, |dE| PC and the FP64 norms that amplify a relative
            # perturbation x of the Koopmans matrices into |dE| ~ norm * |d| * x
            "per_class": [{"class": CLASS_NAMES[k], "dE_sc": float(d[0, k, 1]), "dE_pc": float(d[0, k, 3]),
                           "norm_sc": float(ref[0, k, 0]), "norm_pc": float(ref[0, k, 2])} for k in range(8)],
            "contraction_ms": tmx, "contraction_speedup_vs_fp64": (t64 / tmx) if (t64 and tmx) else None}
        forms[name]["within_1uEh"] = bool(max(d[:, :, 1].max(), d[:, :, 3].max()) <= TOL_EH)
    if "fp64" in forms:
        forms["fp64"]["within_1uEh"] = True
    ok = [k for k in forms if forms[k]["within_1uEh"]]
    head = max(ok, key=lambda k: forms[k]["value"]) if ok else list(forms)[0]
    F = forms[head]
    esize = 8 if head == "fp64" else 4
    peak_c = PEAK[head]
    # rooflines: the K' reduction of the (+-1)' chain (band or tridiagonal form) and the bamat contraction passes
    rl = []
    Nk = n ** 3
    if F["kred_ms"] > 0 and F["kred_form"] == 2:
        fb = band_flops(Nk, 2) * M
        ach = fb / (F["kred_ms"] * 1e-3) / 1e12
        rl.append({"kernel": "K' band reduction (CholeskyQR2 panels, block-reflector GEMMs, two problems batched)",
                   "bound": "mfma", "achieved": ach, "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                   "frac": ach / (PEAK["fp64"] / 1e12), "traffic": None, "device_ms_per_eval": F["kred_ms"],
                   "launches_per_eval": 15 * (Nk // 64) * M,
                   "algorithmic": f"6*64*m^2 + 10*64^2*m per panel of m trailing rows, two problems of order {Nk}: "
                                  f"{fb:.3e} flop per evaluation"})
    elif F["kred_ms"] > 0:
        by = trd_bytes(Nk, 2) * M
        ach = by / (F["kred_ms"] * 1e-3) / 1e9
        rl.append({"kernel": "K' tridiagonal reduction (k_trd_col, one launch per column)", "bound": "hbm",
                   "achieved": ach, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / (HBM_PEAK / 1e9),
                   "traffic": measured_traffic("trd_col", args.config, head),
                   "device_ms_per_eval": F["kred_ms"], "launches_per_eval": (Nk - 1) * M,
                   "algorithmic": f"trailing blocks read and written once per launch, two problems of order {Nk}: "
                                  f"{by:.3e} B per evaluation"})
    t_contr = F["phases"]["E4 contractions"] * 1e-3
    fl, byc = contraction_flops(n, M), contraction_bytes(n, M, esize)
    achc = fl / t_contr / 1e12 if t_contr > 0 else 0.0
    rl.append({"kernel": "bamat 4-RDM contractions (k_e4_pass_a + k_e4_pass_b)", "bound": "mfma", "achieved": achc,
               "peak": peak_c / 1e12, "unit": "TFLOP/s", "frac": achc / (peak_c / 1e12),
               "traffic": measured_traffic("contractions", args.config, head), "device_ms_per_eval": t_contr * 1e3,
               "launches_per_eval": 2 * M,
               "algorithmic": f"8*n^9 flop per state (four contractions of E4 with the active ERIs) = {fl:.3e} "
                              f"flop per evaluation, against the {head.upper()} MFMA peak",
               "bytes": {"two_pass_algorithmic": byc, "achieved_GBps": byc / t_contr / 1e9 if t_contr > 0 else 0.0,
                         "note": "E4 read once per pass + outputs; E4 is materialised, so these bytes are the "
                                 "unfused design's own traffic, not required work"}})
    t_e4 = F["phases"]["E4 build"] * 1e-3
    by4 = e4_build_bytes(n, M, esize)
    ach4 = by4 / t_e4 / 1e9 if t_e4 > 0 else 0.0
    rl.append({"kernel": "E4 build (k_e4_d3t + k_build_e4, pair multisets)", "bound": "hbm", "achieved": ach4,
               "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach4 / (HBM_PEAK / 1e9),
               "traffic": measured_traffic("e4_build", args.config, head),
               "device_ms_per_eval": t_e4 * 1e3, "launches_per_eval": 2 * M,
               "algorithmic": f"E4 written once + packed Gamma4 and 3-RDM read once = {by4:.3e} B per evaluation"})
    # the headline line is the hot path's own MFMA kernel pair (the 4-RDM contractions): a single kernel family
    # whose HIP-event time rocprofv3's per-kernel averages reproduce; the K' reduction (hundreds of small launches
    # under contention) and the E4 build are reported beside it
    dom = next(r for r in rl if r["kernel"].startswith("bamat 4-RDM"))
    res = F["res"]
    out = {
        "metric": METRIC, "value": F["value"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": F["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[head],
        "data": "synthetic (seeded random CI vector -> RDMs built on the GPU; seeded Cholesky ERIs); each of the "
                "concurrent requests runs on its own context with its own device-resident copy of the inputs",
        "config": {"workload": desc, "nact": n, "ncore": nc, "nvirt": nv, "roots": M, "nchovec": P.B.shape[0],
                   "determinants_kept": keep, "precision": head, "concurrency": S},
        "latency_ms": F["latency_ms"],
        "phases_ms": F["phases"],
        "roofline": dom,
        "rooflines": rl,
        "forms": {k: {"value": v["value"], "ms_per_step": v["ms_per_step"], "latency_ms": v["latency_ms"],
                      "within_1uEh_of_fp64": v["within_1uEh"], **({"vs_fp64": v["vs_fp64"]} if "vs_fp64" in v else {})}
                  for k, v in forms.items()},
        "energies": {"e_sc_total": float(res[0, :, 1].sum()), "e_pc_total": float(res[0, :, 3].sum())},
        "setup": {"rdm_build_and_upload_s_per_context": t_setup},
    }
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle on a bounded sample ...", file=sys.stderr, flush=True)
        host_rdms(P, evs[0])
        pairs, frac = CPU_SAMPLE[args.config]
        out["cpu_baseline"] = cpu_baseline(P, pairs, frac)
    for ev in evs:
        ev.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
