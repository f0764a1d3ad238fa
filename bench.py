"""Benchmark: Monte Carlo Fisher realizations per second for the power spectrum + bispectrum at 256^3.

The metric compares two legs measured in the same run.  The timed (reported) leg: one step = R
Monte Carlo realizations (--realizations, default 4) of both Fisher matrices on one GPU, the R Pk
seeds sharing every pass and the Bk Fisher contraction on the matrix cores in split BF16
(pb3d_fisher_contract, accumulating into the FP64 sum), the Bk Q-map chains dealt to --streams HIP
streams (default 4); the plain leg: one realization at a time on one stream with the Bk contraction as
an FP32 GEMM (hipBLASLt).  vs_baseline = timed / plain.  Each realization:
  * PSpec MC Fisher: 256^3, box 1000 Mpc/h, local (Yamamoto) sightline, lmax=4, 30 bins 0.01-0.30 h/Mpc;
  * BSpec MC Fisher: same grid, lmax=2, 40 bins 0.01-0.40 h/Mpc (12700 triangle bins);
  * mask: shell 300 <= r <= 480 Mpc/h minus 200 random holes (r = 10-30 Mpc/h), smoothed 5 Mpc/h;
  * GRFs drawn on the GPU (power law P(k) = 2e4 (k/0.1)^-1.5), Fisher rows accumulated in FP64.
Prints one JSON line with the throughput, the roofline of the dominant kernel (HIP events around
every launch of one single-stream realization of the timed path, run right after the timed steps:
under the multi-stream schedule launches overlap, so their durations would not time one kernel) and a
CPU baseline (the FP64 oracle's unit operations timed on this host and scaled by the reference's exact
per-realization operation counts).

usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--n 256] [--realizations R] [--plain-steps P]
                      [--streams S] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fisher MC realizations/s, Pk+Bk 256³: batched/split-bf16 vs plain FP32"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled x2.000 / x1.000
# by a 1 GiB device copy), averaged over the launches of one realization of this workload, for the
# kernel they were taken on (profiles/hbm_traffic_pk_bk_mc_fisher_256.md); other kernels report null.
# The fused rows' algorithmic bytes count the mask once per map; its rows are shared by every map of a
# launch and the XCD-aware block order lets L2 serve them, so the measured traffic is lower.
MEASURED_TRAFFIC = {"pointing_rows<128>": 5.29e9}
# wave64 VALU instructions per algorithmic byte, rocprofv3 --pmc SQ_INSTS_VALU summed over the launches
# of one realization of this workload divided by their algorithmic bytes (the sampler's, period 1)
# (profiles/pointing_rows_sq_counters_pk_bk_mc_fisher_256.md)
MEASURED_VALU_PER_BYTE = {"pointing_rows<128>": 0.1139}
SIMDS = 256 * 4          # MI355X: 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9         # peak engine clock
KIND_NAMES = {0: "r2c_rows", 1: "c2r_rows", 2: "fft_cols (plain C2R)", 3: "fft_cols (C2R first pass, fused prologue)",
              4: "fft_cols (R2C last pass, fused epilogue)", 5: "pointing_rows (fused Q-map rows)",
              6: "fft_cols (plain R2C)"}


def kind_kernel(kind, n):
    """Kernel template of a sampled pass kind at grid n^3 (the name rocprofv3 reports)."""
    return {0: "r2c_rows<%d>" % (n // 2), 1: "c2r_rows<%d>" % (n // 2), 2: "fft_cols<%d, 1, 0>" % n,
            3: "fft_cols<%d, 1, 1>" % n, 4: "fft_cols<%d, -1, 2>" % n, 5: "pointing_rows<%d>" % (n // 2),
            6: "fft_cols<%d, -1, 0>" % n}[kind]


def shell_mask(n, L, center, seed=1, r_lo=300., r_hi=480., nholes=200, smooth=5.):
    """1 inside r_lo <= r <= r_hi (observer at the origin) minus nholes spheres of radius U[10,30]
    centred uniformly in the shell, smoothed with a Gaussian of width `smooth` (via FFT)."""
    ax = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + center[i] for i in range(3)]
    X = np.meshgrid(*ax, indexing='ij', sparse=True)
    r2 = X[0] ** 2 + X[1] ** 2 + X[2] ** 2
    m = ((r2 >= r_lo ** 2) & (r2 <= r_hi ** 2)).astype(np.float32)
    rng = np.random.RandomState(seed)
    dx = L / n
    for _ in range(nholes):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c = u * rng.uniform(r_lo, r_hi)
        rad = rng.uniform(10., 30.)
        sl = []
        for i in range(3):
            idx = np.nonzero(np.abs(ax[i] - c[i]) <= rad + dx)[0]
            sl.append(idx)
        if any(len(s) == 0 for s in sl):
            continue
        sub = np.ix_(*sl)
        d2 = (ax[0][sl[0]][:, None, None] - c[0]) ** 2 + (ax[1][sl[1]][None, :, None] - c[1]) ** 2 + \
             (ax[2][sl[2]][None, None, :] - c[2]) ** 2
        blk = m[sub]
        blk[d2 < rad ** 2] = 0.
        m[sub] = blk
    k = [np.fft.fftfreq(n, L / n) * 2 * np.pi] * 2 + [np.fft.rfftfreq(n, L / n) * 2 * np.pi]
    K = np.meshgrid(*k, indexing='ij', sparse=True)
    g = np.exp(-0.5 * smooth ** 2 * (K[0] ** 2 + K[1] ** 2 + K[2] ** 2))
    return np.fft.irfftn(np.fft.rfftn(m) * g, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32)


def powerlaw_pk(kmax):
    k = np.linspace(0., kmax, 40001)
    P = np.empty_like(k)
    P[1:] = 2e4 * (k[1:] / 0.1) ** -1.5
    P[0] = P[1]
    return [k, P]


def cpu_baseline(n, pk_bins, bk_bins, lmax_pk, lmax_bk, N_bins_bk):
    """FP64 oracle unit operations at n^3 on this host, scaled by the reference's per-realization
    operation counts: FFTs (scipy.fft, all cores) + the reference's per-bin full-grid sweeps
    (oracle/cpu_kernels.c, OpenMP).  Element-wise passes are not counted (a lower bound on the
    reference's time)."""
    import scipy.fft
    from oracle import polybin_oracle as O
    from oracle import build_ref
    cores = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS"):
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    lib_path = os.path.join(ROOT, "oracle", "_cpu", "libpb3d_cpu.so")
    if not os.path.exists(lib_path):
        build_ref.build_cpu()
    lib = ctypes.CDLL(lib_path)
    dp = ctypes.POINTER(ctypes.c_double)
    t_start = time.perf_counter()
    rng = np.random.RandomState(0)
    x = rng.randn(n, n, n)

    reps = {}

    def med(f, budget=2.5, name=None):
        """median time of f over repetitions filling about `budget` seconds (at least 3)"""
        f()
        ts = []
        while len(ts) < 3 or sum(ts) < budget:
            a = time.perf_counter()
            f()
            ts.append(time.perf_counter() - a)
        if name:
            reps[name] = len(ts)
        return float(np.median(ts))

    X = scipy.fft.rfftn(x, workers=cores)
    t_f = med(lambda: scipy.fft.rfftn(x, workers=cores), name='rfftn')
    t_i = med(lambda: scipy.fft.irfftn(X, s=(n, n, n), workers=cores), name='irfftn')
    kF = 2 * np.pi / 1000.
    kx = np.fft.fftshift(np.arange(-n // 2, n // 2)) * kF
    kz = np.arange(n // 2 + 1) * kF
    modk = np.ascontiguousarray(np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2))
    degk = np.ascontiguousarray(np.broadcast_to(1. + (kz > 0), modk.shape)).astype(np.float64)
    H = modk.size
    nout = max(4, cores)
    arr = np.full((nout, H), 1. + 0.5j, dtype=np.complex128)
    s1 = np.ascontiguousarray(X)
    edges = np.ascontiguousarray(pk_bins, dtype=np.float64)
    out = np.zeros(max(nout, len(edges)))
    lib.cpu_cross_integrate.argtypes = [dp, dp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, dp, ctypes.c_int, dp]
    lib.cpu_integrate_row.argtypes = [dp, dp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, dp]
    P = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    t_cross = med(lambda: lib.cpu_cross_integrate(P(modk), P(degk), s1.ctypes.data, s1.ctypes.data, H, P(edges),
                                                  len(edges) - 1, P(out)), name='cross_integrate')
    t_row = med(lambda: lib.cpu_integrate_row(P(modk), P(degk), s1.ctypes.data, arr.ctypes.data, H, nout, 0.1, 0.11,
                                              P(out)), name='integrate_row')
    t_visit = t_row / nout
    Nk_pk = len(pk_bins) - 1
    Nl_pk = lmax_pk // 2 + 1
    pk_f, pk_r = O.pk_fisher_fft_count(Nk_pk, lmax_pk, True, False)
    ab = O.triangle_bins(bk_bins, bk_bins, lmax_bk)
    bk_f, bk_r = O.bk_fisher_fft_count(ab, len(bk_bins) - 1, lmax_bk, True, False)
    # Pk rows: Nl*Nk Q maps x Nl cross_integrate calls (pspec.py:394-413); Bk rows: one
    # integrate_row over all N_bins Q maps per row (bspec.py:597)
    T = (pk_f + bk_f) * t_f + (pk_r + bk_r) * t_i + Nl_pk * Nk_pk * Nl_pk * t_cross + N_bins_bk ** 2 * t_visit
    sample = ("extrapolated: median times at %d^3 on %d host threads, each op repeated to fill ~2.5 s "
              "(rfftn x%d: %.3f s, irfftn x%d: %.3f s, scipy.fft FP64; one %d-bin cross_integrate sweep x%d: %.3f s; "
              "integrate_row over %d maps x%d: %.2e s per map-sweep; C/OpenMP restatement of the reference loops), "
              "scaled by the reference's %d+%d FFTs, %d cross_integrate calls and %d^2 integrate_row map-sweeps per "
              "realization; element-wise passes not counted; %.1f s of CPU work in this run"
              % (n, cores, reps['rfftn'], t_f, reps['irfftn'], t_i, Nk_pk, reps['cross_integrate'], t_cross, nout,
                 reps['integrate_row'], t_visit, pk_f + bk_f, pk_r + bk_r, Nl_pk * Nk_pk * Nl_pk, N_bins_bk,
                 time.perf_counter() - t_start))
    return {"value": 1.0 / T, "unit": "realizations/s", "cores": cores, "kind": "port", "sample": sample,
            "seconds_per_realization": T}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    # sampling period of the launch timer over the roofline leg (1: every launch of that realization;
    # a larger period should be odd: launches of one kind come in patterns of even length)
    ap.add_argument("--prof-period", type=int, default=1)
    ap.add_argument("--workload", default="mc_fisher", choices=["mc_fisher", "bk_ideal", "pk_mc_fisher", "extras"],
                    help="mc_fisher: the Pk+Bk MC Fisher metric (default); bk_ideal: Bk_ideal evaluations "
                         "and the MFMA triplet contraction (--n, --bk-bins, --lmax, --precision); pk_mc_fisher: "
                         "PSpec MC Fisher realizations at 128^3 and 256^3 (--realizations per call); extras: "
                         "seconds per call of the members around the MC Fisher (shot noise, linear term, "
                         "integral constraints, covariance and theory matrices) at 256^3")
    ap.add_argument("--realizations", type=int, default=None,
                    help="realizations per step: mc_fisher (default 4; the Pk seeds of a step share every pass) "
                         "and pk_mc_fisher (default 1, seeds per compute_fisher batch)")
    ap.add_argument("--graph", action="store_true",
                    help="mc_fisher, pk_mc_fisher: replay every pass as a captured HIP graph (FisherGraph)")
    ap.add_argument("--grids", default="128,256",
                    help="pk_mc_fisher: comma-separated grid sizes")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="mc_fisher: realizations timed on the plain FP32 single-realization path (0 = skip)")
    ap.add_argument("--mixed", action="store_true",
                    help="pk_mc_fisher: issue the 128^3 and 256^3 realizations together, on a stream per grid "
                         "(compute_fisher_mixed), instead of one grid after the other")
    ap.add_argument("--bk-bins", type=int, default=16)
    ap.add_argument("--bk-kmax", type=float, default=0.17)
    ap.add_argument("--lmax", type=int, default=0)
    ap.add_argument("--precision", default="bf16x6", choices=["fp32", "bf16x3", "bf16x6"])
    ap.add_argument("--streams", type=int, default=4,
                    help="mc_fisher: HIP streams the timed leg deals the Bk Q-map chains to (BSpec chain_streams; "
                         "the plain and roofline legs use one)")
    ap.add_argument("--rows-mode", type=int, default=1, choices=[0, 1],
                    help="fused Q-map rows: 1 = harmonic sub-maps combined per row before the row transforms "
                         "(default), 0 = one row transform per sub-map (both legs)")
    args = ap.parse_args()
    if args.realizations is None:
        args.realizations = 4 if args.workload == "mc_fisher" else 1
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the Fisher realizations of one device and has no "
              "multi-GPU path (realizations would shard across GPUs with one FP64 all-reduce of the Fisher "
              "partials); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import polybin3d_amd as pb
    from polybin3d_amd import _capi as C
    C.check(min(0, C.lib().pb3d_pointing_rows_mode(args.rows_mode)))
    if args.workload == "bk_ideal":
        return bk_ideal_workload(args, torch, pb)
    if args.workload == "pk_mc_fisher":
        return pk_fisher_workload(args, torch, pb)
    if args.workload == "extras":
        return extras_workload(args, torch, pb)

    n, L, center = args.n, 1000., [10., 10., 10.]
    kny = n * np.pi / L
    pk_bins = np.linspace(0.01, min(0.30, 0.7 * kny), 31)
    bk_bins = np.linspace(0.01, min(0.40, 0.99 * kny), 41)
    R = max(1, args.realizations)
    t_setup = time.perf_counter()
    base = pb.PolyBin3D(L, [n] * 3, Pk=powerlaw_pk(2 * kny * 1.8), boxcenter=center, sightline='local', rng='device')
    mask = shell_mask(n, L, center)
    ps = pb.PSpec(base, pk_bins, lmax=4, mask=mask)
    bs = pb.BSpec(base, bk_bins, lmax=2, mask=mask, precision=args.precision, chain_streams=args.streams)
    bs._compute_symmetry_factor()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    Fp = torch.zeros((ps.N_bins, ps.N_bins), dtype=torch.float64, device=base.device)
    Fb = torch.zeros((bs.N_bins, bs.N_bins), dtype=torch.float64, device=base.device)

    def plain_step(seed):
        """The plain FP32 single-realization path: one Pk seed, then one Bk seed whose Fisher
        contraction is an FP32 GEMM (hipBLASLt), each Fisher formed on its own and then accumulated."""
        prec, cs = bs.precision, bs.chain_streams
        bs.precision, bs.chain_streams = 'fp32', 1
        Fp.add_(ps._fisher_contribution_dev(seed))
        Fb.add_(bs._fisher_contribution_dev(seed))
        bs.precision, bs.chain_streams = prec, cs

    def roofline_step(seed):
        """One realization of the timed path with every chain on the current stream, so that each
        sampled launch runs alone and its duration times that kernel."""
        cs = bs.chain_streams
        bs.chain_streams = 1
        Fp.add_(ps._fisher_contribution_dev(seed))
        bs._fisher_contribution_dev(seed, out=Fb)
        bs.chain_streams = cs

    graphs = None

    def batched_step(s0):
        """R realizations: the R Pk seeds share every pass (batched), the Bk seeds run back to back with
        the split-BF16 MFMA contraction accumulating straight into the FP64 Fisher sum (each on a stream
        of its own, beside the next seed's maps and tables); with --graph each of the two passes is one
        HIP-graph replay."""
        if graphs is not None:
            graphs[0].replay(s0)
            graphs[1].replay(s0)
            return
        seeds = list(range(s0, s0 + R))
        Fp.add_(ps._fisher_contributions_dev(seeds) if R > 1 else ps._fisher_contribution_dev(seeds[0]))
        bs._fisher_contributions_dev(seeds, Fb)

    # leg 1: the plain FP32 single-realization path (the metric's comparison point)
    plain = None
    if args.plain_steps > 0:
        plain_step(10 ** 7)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(args.plain_steps):
            plain_step(2 * 10 ** 7 + s)
        torch.cuda.synchronize()
        dtp = time.perf_counter() - t0
        plain = {"value": round(args.plain_steps / dtp, 6), "unit": "realizations/s", "steps": args.plain_steps,
                 "ms_per_realization": round(1e3 * dtp / args.plain_steps, 2),
                 "path": "one Pk seed per pass, Bk Fisher contraction as FP32 GEMM (hipBLASLt), no batching"}
        print("plain FP32 leg: %.1f ms per realization" % (1e3 * dtp / args.plain_steps), file=sys.stderr, flush=True)
    # leg 2 (timed): the batched / split-BF16 path
    if args.graph:
        t = time.perf_counter()
        graphs = (pb.FisherGraph(ps, Fp, seeds_per_replay=R), pb.FisherGraph(bs, Fb, seeds_per_replay=R))
        torch.cuda.synchronize()
        print("captured the Pk and Bk passes (%.1f s)" % (time.perf_counter() - t), file=sys.stderr, flush=True)
    for w in range(args.warmup):
        batched_step(10 ** 6 + w * R)
        print("warmup step %d done" % w, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        batched_step(s * R)
        torch.cuda.synchronize()
        print("timed step %d done (%.2f s)" % (s, time.perf_counter() - t0), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # leg 3: the launch sampler (HIP events on each launch's stream) over one single-stream realization
    # of the timed path: with the chains dealt to several streams, launches of different chains overlap
    # and a launch's duration no longer times one kernel
    C.check(C.lib().pb3d_prof_enable(args.prof_period))
    roofline_step(3 * 10 ** 7)
    torch.cuda.synchronize()
    print("roofline leg done", file=sys.stderr, flush=True)
    kinds = {}
    for kind in sorted(KIND_NAMES):
        cnt, ms, by = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        C.check(C.lib().pb3d_prof_read(kind, ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(by)))
        if cnt.value:
            kinds[kind] = (cnt.value, ms.value, by.value)
    C.lib().pb3d_prof_enable(0)
    dom = max(kinds, key=lambda k: kinds[k][1]) if kinds else None
    roof = None
    if dom is not None:
        cnt, ms, by = kinds[dom]
        achieved = by / (ms * 1e-3) / 1e9
        kname = kind_kernel(dom, n)
        hbm_frac = achieved / HBM_PEAK_GBS
        vpb = MEASURED_VALU_PER_BYTE.get(kname)
        valu_frac = None
        if vpb is not None:
            # VALU-issue fraction: wave64 VALU instructions of the launch (per-byte count x its algorithmic
            # bytes) x 4 cycles, over the SIMD cycles of the measured launch duration
            valu_frac = vpb * (by / cnt) * 4. / (SIMDS * CLOCK_HZ * (ms / cnt) * 1e-3)
        bound = "valu" if (valu_frac is not None and valu_frac > hbm_frac) else "hbm"
        roof = {"bound": bound, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm_frac, 4), "traffic": MEASURED_TRAFFIC.get(kname),
                "valu_issue_frac": None if valu_frac is None else round(valu_frac, 3),
                "bound_from": "rocprofv3 SQ_INSTS_VALU and FETCH_SIZE/WRITE_SIZE of %s on this workload "
                              "(profiles/) with this run's launch durations" % kname,
                "kernel": kname, "pass": KIND_NAMES[dom],
                "measured_on": "one single-stream realization of the timed path, right after the timed steps",
                "avg_launch_ms": round(ms / cnt, 4), "bytes_per_launch": by / cnt,
                "sampled_launches": cnt, "share_of_sampled_fft_time": round(ms / sum(v[1] for v in kinds.values()), 3),
                "passes": {KIND_NAMES[k]: {"sampled_launches": v[0], "avg_launch_ms": round(v[1] / v[0], 4),
                                           "GB_per_s": round(v[2] / (v[1] * 1e-3) / 1e9, 1)}
                           for k, v in sorted(kinds.items())}}
    value = args.steps * R / dt
    res = {"metric": METRIC, "value": round(value, 6), "unit": "realizations/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None if plain is None else round(value / plain["value"], 4),
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "pk_bk_mc_fisher_%d" % n, "grid": [n] * 3, "boxsize": L, "sightline": "local",
                      "pk": {"lmax": 4, "nbins": len(pk_bins) - 1, "kmax": float(pk_bins[-1])},
                      "bk": {"lmax": 2, "nbins": len(bk_bins) - 1, "kmax": float(bk_bins[-1]), "N_bins": bs.N_bins},
                      "mask": "shell 300-480 Mpc/h, 200 holes, 5 Mpc/h smoothing", "rng": "device",
                      "realizations_per_step": R, "pk_seeds_per_pass": R, "hip_graph": bool(args.graph),
                      "pointing_rows_mode": args.rows_mode, "bk_chain_streams": args.streams,
                      "bk_chain_streams_used": getattr(bs, 'streams_used', 1),
                      "bk_fisher_contraction": args.precision + " MFMA (split BF16, FP32 runs, FP64 store)",
                      "fisher_accumulation": "fp64", "setup_s": round(t_setup, 1)},
           "vs_baseline_is": "value / plain FP32 single-realization leg measured in this run (the metric's "
                             "comparison; no published reference throughput exists)",
           "plain_fp32_leg": plain,
           "roofline": roof}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(n, pk_bins, bk_bins, 4, 2, bs.N_bins)
    print(json.dumps(res), flush=True)


def pk_fisher_workload(args, torch, pb):
    """PSpec MC Fisher (local sightline, lmax 4, 30 bins 0.01..0.30, shell mask, device GRFs, FP64
    accumulation) on a 128^3 and a 256^3 grid.  One step = args.realizations realizations on each grid;
    value = realizations per second over both grids."""
    L, center = 1000., [10., 10., 10.]
    grids = tuple(int(v) for v in args.grids.split(','))
    specs = {}
    for n in grids:
        b = pb.PolyBin3D(L, [n] * 3, Pk=powerlaw_pk(2 * n * np.pi / L * 1.8), boxcenter=center, sightline='local',
                         rng='device')
        # 30 bins to 0.30 h/Mpc, or to 0.9 k_Nyquist on grids too small for that
        specs[n] = pb.PSpec(b, np.linspace(0.01, min(0.30, 0.9 * n * np.pi / L), 31), lmax=4,
                            mask=shell_mask(n, L, center))
    R = max(1, args.realizations)
    F = {n: torch.zeros((specs[n].N_bins,) * 2, dtype=torch.float64, device=specs[n].base.device) for n in grids}
    per = {n: 0. for n in grids}

    streams = {n: torch.cuda.Stream() for n in grids}  # default priority
    graphs = {n: pb.FisherGraph(specs[n], F[n], seeds_per_replay=R) for n in grids} if args.graph else None

    def step(seed0, timed):
        seeds = list(range(seed0, seed0 + R))
        if graphs is not None:
            for n in grids:
                t = time.perf_counter()
                graphs[n].replay(seed0)
                torch.cuda.synchronize()
                if timed:
                    per[n] += time.perf_counter() - t
            return
        if args.mixed:
            # every grid's realizations on its own stream (compute_fisher_mixed's schedule)
            main = torch.cuda.current_stream()
            for n in grids:
                streams[n].wait_stream(main)
                with torch.cuda.stream(streams[n]):
                    F[n].add_(specs[n]._fisher_contributions_dev(seeds))
            for n in grids:
                main.wait_stream(streams[n])
            torch.cuda.synchronize()
            return
        for n in grids:
            t = time.perf_counter()
            F[n].add_(specs[n]._fisher_contributions_dev(seeds))
            torch.cuda.synchronize()
            if timed:
                per[n] += time.perf_counter() - t
    for w in range(args.warmup):
        step(10 ** 6 + w * R, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(s * R, True)
    dt = time.perf_counter() - t0
    nreal = args.steps * R * len(grids)
    res = {"metric": "Pk MC Fisher realizations/s (%s)" % " + ".join("%d^3" % n for n in grids),
           "value": round(nreal / dt, 4),
           "unit": "realizations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(1e3 * dt / args.steps, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "pk_mc_fisher_" + "_".join(str(n) for n in grids), "grids": list(grids),
                      "boxsize": L, "sightline": "local", "lmax": 4, "nbins": 30, "kmax": 0.3, "mask": "shell",
                      "realizations_per_batch": R, "hip_graph": bool(args.graph), "fisher_accumulation": "fp64",
                      "schedule": "mixed streams" if args.mixed else "grid by grid"}}
    if not args.mixed:
        res["ms_per_realization"] = {str(n): round(1e3 * per[n] / (args.steps * R), 2) for n in grids}
    print(json.dumps(res), flush=True)


def extras_workload(args, torch, pb):
    """Seconds per call (median of args.steps after args.warmup) of the estimator members around the
    MC Fisher, on the metric's 256^3 configuration (local sightline, shell mask, device GRFs):
    PSpec (lmax 4, 30 bins to 0.30): shot-noise numerator, MC Fisher with the global and the radial
    integral constraint, covariance contribution (P_0, P_2), theory contribution (60 theory bins,
    lmax_theory 4, wide-angle); BSpec (lmax 2, 40 bins to 0.40, 12 700 triangles): numerator with the
    linear term of 4 sims, shot-noise numerator (cubic and quadratic terms); BSpec covariance on a
    smaller case (64^3, 8 bins to 0.16, lmax 2), whose Q maps pass through the filter one at a time."""
    n, L, center = args.n, 1000., [10., 10., 10.]
    kw = dict(boxcenter=center, sightline='local', rng='device')
    b = pb.PolyBin3D(L, [n] * 3, Pk=powerlaw_pk(2 * n * np.pi / L * 1.8), **kw)
    mask = shell_mask(n, L, center)
    kk = np.linspace(0., 2. * n * np.pi / L, 4000)
    Pc = [kk, 2e4 / (1 + (kk / 0.04) ** 2), 5e3 / (1 + (kk / 0.05) ** 2)]
    res = {}

    def timeit(name, f):
        for w in range(args.warmup):
            f(10 ** 6 + w)
        ts = []
        for s in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            f(s)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        res[name] = round(float(np.median(ts)), 4)
        print(json.dumps({name: res[name]}), flush=True)

    pk_bins = np.linspace(0.01, 0.30, 31)
    ps = pb.PSpec(b, pk_bins, lmax=4, mask=mask, mask_shot=mask)
    timeit("pk_shot_contribution", lambda s: ps.compute_shot_contribution(s))
    timeit("pk_covariance_contribution", lambda s: ps.compute_covariance_contribution(s, Pc))
    timeit("pk_theory_contribution_wideangle", lambda s: ps.compute_theory_contribution(
        s, np.linspace(0.01, 0.30, 61), lmax_theory=4, include_wideangle=True))
    x = np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + center[0]
    X = np.meshgrid(x, x, x, indexing='ij')
    modr = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    del X
    pg = pb.PSpec(b, pk_bins, lmax=4, mask=mask, add_GIC=True, mask_IC=mask)
    timeit("pk_mc_fisher_gic", lambda s: pg._fisher_contribution_dev(s))
    pr = pb.PSpec(b, pk_bins, lmax=4, mask=mask, add_RIC=True, mask_IC=mask,
                  radial_bins_RIC=np.linspace(modr[mask > 0.5].min(), modr[mask > 0.5].max() + 1., 9))
    timeit("pk_mc_fisher_ric", lambda s: pr._fisher_contribution_dev(s))
    del ps, pg, pr
    bs = pb.BSpec(b, np.linspace(0.01, 0.40, 41), lmax=2, mask=mask, mask_shot=mask)
    data = b.generate_data(seed=5, output_type='real') * b.as_real(mask)
    bs.generate_sims(4)
    timeit("bk_numerator_linear_term_4sims", lambda s: bs.Bk_numerator(data))
    timeit("bk_shot_contribution", lambda s: bs.compute_shot_contribution(s, data=data))
    del bs
    nb = 64
    b64 = pb.PolyBin3D(L, [nb] * 3, Pk=powerlaw_pk(2 * nb * np.pi / L * 1.8), **kw)
    m64 = shell_mask(nb, L, center)
    bc = pb.BSpec(b64, np.linspace(0.01, 0.16, 9), lmax=2, mask=m64, mask_shot=m64)
    timeit("bk_covariance_contribution_64cubed_%dbins" % bc.N_bins, lambda s: bc.compute_covariance_contribution(s, Pc))
    out = {"metric": "seconds per call of the members around the MC Fisher", "unit": "s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "data": "synthetic",
           "dtype": "f32", "config": {"workload": "extras_256", "grid": [n] * 3, "boxsize": L, "sightline": "local",
                                      "mask": "shell", "rng": "device"},
           "seconds": res}
    print(json.dumps(out), flush=True)


def bk_ideal_workload(args, torch, pb):
    """Bk_ideal on a quadratic non-Gaussian field (delta_G + (delta_G^2 - <delta_G^2>)/2, power-law GRF),
    global sightline, linear bins 0.01..bk_kmax.  One step = one Bk_ideal evaluation (1 R2C, Nl*Nk
    filtered C2Rs, the triplet contraction, the cached inverse ideal Fisher).  The triplet kernel is
    timed with HIP events on the launch stream; its bf16 MFMA rate counts the split products executed
    (3 or 6 per multiply-add) over the dense T[a][p] block."""
    n, L = args.n, 1000.
    edges = np.linspace(0.01, args.bk_kmax, args.bk_bins + 1)
    b = pb.PolyBin3D(L, [n] * 3, Pk=powerlaw_pk(2 * n * np.pi / L * 1.8), rng='device')
    bs = pb.BSpec(b, edges, lmax=args.lmax, precision=args.precision)
    dG = b.generate_data(seed=0, output_type='real')
    field = (dG + 0.5 * (dG ** 2 - (dG ** 2).mean())).contiguous()
    del dG
    bs.Bk_ideal(field)  # ideal Fisher and symmetry factors (data independent), first call
    orig = bs._triplet
    trip_ms, trip_calls = [], []

    def timed_triplet(stack, i1, i2, i3):
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        r = orig(stack, i1, i2, i3)
        e1.record(s)
        e1.synchronize()
        trip_ms.append(e0.elapsed_time(e1))
        i3 = np.asarray(i3)
        pairs = np.unique(np.minimum(i1, i2) * stack.shape[0] + np.maximum(i1, i2))
        trip_calls.append((len(np.unique(i3)), len(pairs), int(np.prod(stack.shape[1:])), len(i3)))
        return r
    bs._triplet = timed_triplet
    for _ in range(args.warmup):
        bs.Bk_ideal(field)
    trip_ms.clear(); trip_calls.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bs.Bk_ideal(field)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nterm = {"fp32": 1, "bf16x3": 3, "bf16x6": 6}[args.precision]
    pad = lambda v, m: (v + m - 1) // m * m  # noqa: E731
    flops = sum(2. * pad(na, 32) * pad(np_, 32) * nv * nterm for na, np_, nv, _ in trip_calls)
    useful = sum(2. * na * np_ * nv for na, np_, nv, _ in trip_calls)
    vec = sum(3. * nt * nv for _, _, nv, nt in trip_calls)  # 2 multiplies + 1 add per triplet and voxel
    ms = sum(trip_ms)
    mfma = args.precision != "fp32"
    res = {"metric": "Bk_ideal evaluations/s", "value": round(args.steps / dt, 4), "unit": "evaluations/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": args.precision, "data": "synthetic",
           "config": {"workload": "bk_ideal_%d" % n, "grid": [n] * 3, "boxsize": L, "sightline": "global",
                      "lmax": args.lmax, "nbins": args.bk_bins, "kmax": args.bk_kmax, "N_bins": bs.N_bins,
                      "triplet_precision": args.precision},
           "triplet": {"ms_per_step": round(ms / args.steps, 3), "share_of_step": round(ms * 1e-3 / dt, 3),
                       "A_maps_pairs_voxels_triplets": trip_calls[0] if trip_calls else None,
                       "fp32_equivalent_TFLOP_s": round(useful / (ms * 1e-3) / 1e12, 2)},
           "roofline": ({"bound": "mfma", "achieved": round(flops / (ms * 1e-3) / 1e12, 2), "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": round(flops / (ms * 1e-3) / 1e12 / 2500.0, 4), "traffic": None,
                         "kernel": "pair_gemm_fwd (bf16 MFMA, %d products per FP32 multiply-add)" % nterm}
                        if mfma else {"bound": "valu", "achieved": round(vec / (ms * 1e-3) / 1e12, 2),
                                      "peak": 157.3, "unit": "TFLOP/s",
                                      "frac": round(vec / (ms * 1e-3) / 1e12 / 157.3, 4), "traffic": None,
                                      "kernel": "triplet_vec_kernel (FP32 products, FP64 sums, vector ALU)"})}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
