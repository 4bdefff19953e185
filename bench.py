#!/usr/bin/env python3
"""Benchmark of the calc + update hot path on one MI355X.

Workload (default): the one-GPU tile -- 51 layers x 1500 x 1800 curvilinear
grid, ensemble m = 96, 7M super-obs (60% SST uniform, 30% SLA along 500 km
tracks, 10% T/S profiles over 51 levels), LOCRAD 250 km, STRIDE 3 (300,000
local analyses), plain FP64 GPU path.  Synthetic data generated on the GPU
(Gaussian random fields, 300 km correlation length).

A step is one pass of the hot path: das_calctransforms (localisation, per-
column transforms, diagnostics) for every stride node, then das_updatefields
over all 103 fields of the tile, with every input already resident in HBM.
`value` is local analyses per second of the calc pass; the update rate
(cell-member updates per second) is reported beside it.

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("enkf-c_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

METRIC = "local analyses/s + cell·member updates/s at m=96; mixed vs FP64 GPU path"
PEAK_FP64 = 78.6e12     # MI355X dense FP64 (matrix and vector), flop/s
PEAK_HBM = 8.0e12       # MI355X HBM3E, B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(workload_tag, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE passes, each scaled
    by the reading of a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "%s_hbm_traffic.md" % workload_tag)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].startswith("`%s`" % kernel):
            try:
                return float(cells[4]) * 1e9
            except ValueError:
                return None
    return None


def cpu_baseline(X, scheme_id, md, sample_target_s=15.0, max_nodes=4000):
    """The reference kernels (oracle/_ref: calc/calcs.c + common/kdtree.c with
    MKL, one thread) or, when they are not built, the oracle restatement, timed
    on a contiguous sample of stride nodes from the middle of the tile."""
    import numpy as np
    import oracle as O
    os.environ.setdefault("MKL_NUM_THREADS", "1")
    kind = "reference" if O.ref_available() else "port"
    try:
        if kind == "reference":
            O.ref()
    except OSError:
        kind = "port"
    ob = O.OracleObs(X.obstypes, X.obs_type, X.obs_lon, X.obs_lat, X.obs_estd, build_trees=(kind == "port"))
    S, s = O.standardise(X.nmem, ob, X.HA, X.innov)
    handle = O.RefTrees(ob) if kind == "reference" else False
    node0 = (X.nj_s // 2) * X.ni_s + X.ni_s // 4
    lon = X.node_lon.ravel()[node0:node0 + max_nodes]
    lat = X.node_lat.ravel()[node0:node0 + max_nodes]
    # calibrate on a few nodes, then time a sample of about sample_target_s
    t = time.perf_counter()
    O.calc_transforms(1, scheme_id, X.nmem, md, 1.0, ob, S, s, lon[:8], lat[:8], use_ref=handle)
    per = (time.perf_counter() - t) / 8
    n = int(max(16, min(max_nodes, sample_target_s / max(per, 1e-6))))
    t = time.perf_counter()
    r = O.calc_transforms(1, scheme_id, X.nmem, md, 1.0, ob, S, s, lon[:n], lat[:n], use_ref=handle)
    dt = time.perf_counter() - t
    pbar = r["stats"]["nlobs_sum"] / max(r["stats"]["ncell"], 1)
    what = ("reference calc/calcs.c + common/kdtree.c (MKL, 1 thread) driven by the restated "
            "das_calctransforms loop" if kind == "reference" else "oracle restatement (plain C, 1 thread)")
    return dict(value=n / dt, unit="local analyses/s", cores=1, kind=kind,
                sample="%d consecutive stride nodes of the middle row band (mean %.0f local obs), %s, %.1f s"
                       % (n, pbar, what, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scheme", choices=["denkf", "etkf"], default="denkf")
    ap.add_argument("--size", choices=["full", "small"], default="full")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "both"], default="both",
                    help="both: the FP64 path (value) and then the mixed path, reported beside it")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        log("bench.py measures one GPU: the path would shard by stride-row bands balanced by local "
            "obs counts, but only the one-GPU run is measured.  Use --gpus 1.")
        sys.exit(2)

    import numpy as np
    import torch
    import enkfg
    import synth_gpu

    if args.size == "full":
        tile = dict(nj=1500, ni=1800, nk=51, nmem=96, n_sst=4_200_000, n_sla=2_100_000, n_prof=6_900)
    else:
        tile = dict(nj=300, ni=360, nk=10, nmem=96, n_sst=170_000, n_sla=84_000, n_prof=280)
    scheme_id = enkfg.SCHEME_DENKF if args.scheme == "denkf" else enkfg.SCHEME_ETKF
    log("generating the synthetic tile %s on the GPU ..." % tile)
    t0 = time.perf_counter()
    X = synth_gpu.make_tile(**tile, seed=0)
    log("  done in %.1f s: %d obs, %d x %d nodes, %d fields" % (time.perf_counter() - t0, X.nobs, X.nj_s, X.ni_s,
                                                                 X.nfield))
    m = X.nmem
    md = m
    n_an = X.nj_s * X.ni_s
    nfield, ncell = X.nfield, X.ncell
    # a band of stride rows in the middle of the tile, for the mixed-versus-FP64 comparison
    jb = X.nj_s // 2
    sample = (X.node_lon[jb:jb + 8], X.node_lat[jb:jb + 8])

    def run(precision):
        """warmup + timed steps of one precision; returns rates, phases and the
        transforms of the comparison band"""
        E = enkfg.Engine(m, mode=enkfg.MODE_ENKF, scheme=scheme_id, precision=precision, pertype_diag=True,
                         keep_fp64=False)
        E.set_obstypes(X.obstypes)
        t0 = time.perf_counter()
        E.set_obs(X.obs_type, X.obs_status, X.obs_lon, X.obs_lat, X.obs_estd)
        t_tree = time.perf_counter() - t0
        E.set_S(X.HA, X.innov, standardised=False)
        log("  [%s] obs + kd-trees set in %.1f s, S uploaded and standardised" % (pname[precision], t_tree))
        for w in range(max(args.warmup, 1)):
            E.calc_transforms(X.node_lon, X.node_lat, stride=X.stride)
            ph = E.timings()
            if w == 0:
                E.ens_attach(X.nlevels, X.stride, 0, 0, X.fields, X.E.data_ptr())
            if w < args.warmup:
                E.ens_update()
                ph[3] = E.timings()[3]
                log("  [%s] warmup %d: phases (ms) loc %.1f gram %.1f factor %.1f update %.1f"
                    % (pname[precision], w, *ph))
        torch.cuda.synchronize()
        E.sync()
        phases = np.zeros(4)
        tcalc = tupd = 0.0
        t0 = time.perf_counter()
        for k in range(args.steps):
            a = time.perf_counter()
            E.calc_transforms(X.node_lon, X.node_lat, stride=X.stride)
            b = time.perf_counter()
            ph = E.timings()
            E.ens_update()
            c = time.perf_counter()
            ph[3] = E.timings()[3]
            phases += np.array(ph)
            tcalc += b - a
            tupd += c - b
            log("  [%s] step %d: calc %.1f ms update %.1f ms (loc %.1f gram %.1f factor %.1f update-kernel %.1f)"
                % (pname[precision], k, 1e3 * (b - a), 1e3 * (c - b), *ph))
        E.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        K = args.steps
        st = E.stats()
        nl = E.launches()
        tuning = E.tuning()   # bucket counts of the timed tile calc
        # transforms of the comparison band (after the timed region)
        E.calc_transforms(*sample, stride=X.stride)
        Tb, wb, _, _ = E.transforms()
        pnb, pdb, _ = E.diag()
        E.close()
        return dict(calc_rate=n_an / (tcalc / K), upd_rate=nfield * ncell * md / (tupd / K), phases=phases / K,
                    tcalc=tcalc / K, tupd=tupd / K, wall=wall, stats=st, launches=nl, t_tree=t_tree,
                    band=(Tb, wb, pdb), tuning=tuning)

    def run_graph(precision):
        """calc passes of the same workload with the Gram/factor batches as
        one captured HIP graph (ENKFG_FLAG_GRAPH): two warmup passes (tuning,
        capture), then timed replays; returns ms per calc pass"""
        E = enkfg.Engine(m, mode=enkfg.MODE_ENKF, scheme=scheme_id, precision=precision, pertype_diag=True,
                         keep_fp64=False, graph=True)
        E.set_obstypes(X.obstypes)
        E.set_obs(X.obs_type, X.obs_status, X.obs_lon, X.obs_lat, X.obs_estd)
        E.set_S(X.HA, X.innov, standardised=False)
        for _ in range(2):
            E.calc_transforms(X.node_lon, X.node_lat, stride=X.stride)
        E.sync()
        t0 = time.perf_counter()
        for k in range(args.steps):
            E.calc_transforms(X.node_lon, X.node_lat, stride=X.stride)
        E.sync()
        ms = 1e3 * (time.perf_counter() - t0) / args.steps
        log("  [%s, graph] calc %.1f ms per pass (%d replays)" % (pname[precision], ms, args.steps))
        E.close()
        return ms

    pname = {enkfg.PREC_FP64: "fp64", enkfg.PREC_MIXED: "mixed"}
    R = run(enkfg.PREC_FP64)
    Rm = run(enkfg.PREC_MIXED) if args.precision == "both" else None
    graph_ms = run_graph(enkfg.PREC_FP64)
    K = args.steps
    phases = R["phases"]
    st = R["stats"]
    calc_rate, upd_rate = R["calc_rate"], R["upd_rate"]
    tag = "tile_fp64_%s" % args.scheme if args.size == "full" else "small_fp64_%s" % args.scheme

    # algorithmic work (per step)
    p_sum = st["nlobs_sum"]
    F_gram = p_sum * (m * (m + 1) + 2 * m)              # S^T S upper triangle (SYRK count) + S^T s
    lev = X.nlevels
    wet = 0
    for f in X.fields:
        thr = 0 if f.get("issurfacevar") else f["level"]
        wet += int((lev > thr).sum())
    B_upd = 4 * m * (2 * wet + (nfield * ncell - wet)) + 4 * n_an * (m * md + m)
    # the factor kernels run on a second stream beside the next batch's Gram,
    # so their summed durations are stretched by that overlap and do not
    # measure their own cost: the dominant kernel is chosen among the others
    kern = {"gram": phases[1], "update": phases[3], "loc": phases[0]}
    dom = max(kern, key=kern.get)
    # the update kernel against HBM, reported whichever kernel dominates
    uname = "update_mfma_kernel<%d, false>" % ((m + 15) // 16)
    uach = B_upd / (phases[3] * 1e-3)
    upd_roof = dict(bound="hbm", achieved=uach / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s", frac=uach / PEAK_HBM,
                    traffic=profiled_traffic(tag, uname), kernel=uname, launches_per_step=1,
                    avg_launch_ms=phases[3],
                    algorithmic="4*m*(2*wet cell-fields + dry cell-fields) + 4*N_an*(m*m_d+m) = %.3e B per launch"
                                % B_upd)
    if dom == "update":
        roof = upd_roof
    elif dom == "gram":
        ach = F_gram / (phases[1] * 1e-3)
        # the item kernel in every tile shape the autotuner chose (per
        # obs-count bucket): the FP64 Gram phase is their launches together
        kname = ("gram_item_kernel<%d, *>" % ((m + 15) // 16) if m <= 96 else "gram_kernel<%d>" % ((m + 15) // 16))
        nl = R["launches"]
        roof = dict(bound="mfma", achieved=ach / 1e12, peak=PEAK_FP64 / 1e12, unit="TFLOP/s", frac=ach / PEAK_FP64,
                    traffic=profiled_traffic(tag, kname), kernel=kname + " (FP64 MFMA)",
                    launches_per_step=nl, avg_launch_ms=phases[1] / max(nl, 1),
                    algorithmic="sum_nodes p*(m(m+1)+2m) = %.3e flop per step over %d launches" % (F_gram, nl))
    else:
        roof = dict(bound="latency", achieved=None, peak=None, unit=None, frac=None, traffic=None,
                    kernel=dom, algorithmic="dominant phase is %s" % dom)

    res = dict(metric=METRIC, value=calc_rate, unit="local analyses/s", n_gpus=1, steps=K, warmup=args.warmup,
               ms_per_step=1e3 * (R["wall"]) / K, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64",
               data="synthetic (GPU-generated Gaussian random fields; no files)",
               config=dict(workload="tile %dx%dx%d, m=%d, %d obs, %s, FP64, stride %d, locrad 250 km, "
                                    "%d local analyses, %d fields" % (X.nk, X.nj, X.ni, m, X.nobs,
                                                                      args.scheme.upper(), X.stride, n_an, nfield),
                           grid=[X.nk, X.nj, X.ni], nmem=m, nobs=int(X.nobs), n_an=n_an, nfield=nfield,
                           precision="fp64", scheme=args.scheme, mean_local_obs=p_sum / n_an),
               update_cell_member_per_s=upd_rate,
               phases_ms=dict(localise=phases[0], gram=phases[1], factor=phases[2], update=phases[3],
                              calc_wall=1e3 * R["tcalc"], update_wall=1e3 * R["tupd"]),
               kd_tree_build_s=R["t_tree"],
               roofline=roof,
               update_roofline=upd_roof,
               calc_buckets=R["tuning"],
               graph=dict(note="the same FP64 calc pass with the Gram/factor batches replayed as one captured "
                               "HIP graph (ENKFG_FLAG_GRAPH)", calc_ms=graph_ms,
                          local_analyses_per_s=n_an / (graph_ms * 1e-3),
                          speedup_vs_stream_launches=1e3 * R["tcalc"] / graph_ms))
    if Rm is not None:
        Tb, wb, pdb = R["band"]
        Tm, wm, pdm = Rm["band"]
        nz = pdb != 0
        F_mix = 3 * p_sum * (m * (m + 1) + 2 * m)       # three split products
        pm = Rm["phases"]
        res["mixed"] = dict(
            precision="S^T S as BF16 hi/lo MFMA (3 products, FP32 accumulation); factorisation FP64; update FP32",
            local_analyses_per_s=Rm["calc_rate"], update_cell_member_per_s=Rm["upd_rate"],
            calc_speedup_vs_fp64=Rm["calc_rate"] / calc_rate,
            step_speedup_vs_fp64=(R["tcalc"] + R["tupd"]) / (Rm["tcalc"] + Rm["tupd"]),
            phases_ms=dict(localise=pm[0], gram=pm[1], factor=pm[2], update=pm[3],
                           calc_wall=1e3 * Rm["tcalc"], update_wall=1e3 * Rm["tupd"]),
            gram_bf16_tflops=F_mix / (pm[1] * 1e-3) / 1e12,
            calc_buckets=Rm["tuning"],
            error_vs_fp64=dict(
                sample="8 stride rows (%d nodes) in the middle of the tile" % (8 * X.ni_s),
                T=float(np.abs(Tm - Tb).max() / np.abs(Tb).max()),
                w=float(np.abs(wm - wb).max() / max(np.abs(wb).max(), 1e-30)),
                dfs=float(np.abs(pdm[nz] - pdb[nz]).max() / np.abs(pdb[nz]).max())))
    if not args.no_cpu_baseline:
        log("CPU baseline ...")
        del X.E
        torch.cuda.empty_cache()
        res["cpu_baseline"] = cpu_baseline(X, scheme_id, md)
        log("  %s" % res["cpu_baseline"])
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
