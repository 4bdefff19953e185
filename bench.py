#!/usr/bin/env python3
"""Benchmark of the calc + update hot path on one MI355X.

Workload (default, --config c2): the one-GPU tile -- 51 layers x 1500 x 1800
curvilinear grid, ensemble m = 96, 7M super-obs (60% SST uniform, 30% SLA
along 500 km tracks, 10% T/S profiles over 51 levels), LOCRAD 250 km,
STRIDE 3 (300,000 local analyses), plain FP64 GPU path, then the same tile
in mixed precision.  --config c4: the hybrid 96 + 48 member problem on a
51 x 1200 x 1200 grid with clustered obs (log-normal local counts), mixed
precision as the value.  Synthetic data generated on the GPU by
libenkfsynth.so (no other GPU library is loaded).

A step is one pass of the hot path: das_calctransforms (localisation, per-
column transforms, diagnostics) for every stride node, then das_updatefields
over all 103 fields of the tile, with every input already resident in HBM.
`value` is local analyses per second of the calc pass; the update rate
(cell-member updates per second) is reported beside it.

After the timed steps: the reference CPU path (calc/calcs.c + common/kdtree.c
with MKL, driven by the restated loop, one thread per host core) on a band of
stride rows, the engine's FP64 and mixed transforms of the same band compared
with it, and the update of the band's model rows by the restated FP32
das_updatefields on the host cores, by the engine from host buffers (PCIe
included) and by the mixed engine, compared with each other.

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("enkf-c_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

METRIC = "local analyses/s + cell·member updates/s at m=96; mixed vs FP64 GPU path"
PEAK_FP64 = 78.6e12     # MI355X dense FP64 (matrix and vector), flop/s
PEAK_HBM = 8.0e12       # MI355X HBM3E, B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """CPU threads for the CPU baseline: the cores this process may use,
    at most 16 (a one-GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


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


def relerr(a, b):
    import numpy as np
    return float(np.abs(a.astype(np.float64) - b).max() / max(float(np.abs(b).max()), 1e-300))


def band_buffers(X, r0, rows):
    """host field buffers [nmem][rows][ni] of model rows r0 .. r0 + rows - 1
    for every field, read from the device ensemble [field][cell][member]"""
    import numpy as np
    m, ni = X.nmem, X.ni
    out = []
    for f in range(X.nfield):
        a = X.E.read((f * X.ncell + r0 * ni) * m, rows * ni * m).reshape(rows * ni, m)
        out.append(np.ascontiguousarray(a.T).reshape(m, rows, ni))
    return out


def cpu_side(X, args, mode, scheme_id, md, engine_kw):
    """The reference CPU path on a band of stride rows (calc) and its model
    rows (update), timed on the host cores, and the engine's results on the
    same band compared with it."""
    import numpy as np
    import enkfg
    import oracle as O
    nth = host_cores()
    kind = "reference" if O.ref_available() else "port"
    try:
        if kind == "reference":
            O.ref()
    except OSError:
        kind = "port"
    use_ref = kind == "reference"
    ob = O.OracleObs(X.obstypes, X.obs_type, X.obs_lon, X.obs_lat, X.obs_estd, build_trees=not use_ref)
    S, s = O.standardise(X.nmem, ob, X.HA, X.innov)
    jj0 = X.nj_s // 2
    lon_all, lat_all = X.node_lon, X.node_lat
    # one set of the reference's kd-trees per thread (its search keeps its
    # results in the tree), built before any timing
    t = time.perf_counter()
    trees = O.ref_trees_parallel(ob, nth) if use_ref else False
    log("CPU baseline: %d sets of kd-trees built in %.1f s" % (nth if use_ref else 0, time.perf_counter() - t))
    # calibrate on 8 nodes of the band's first row, one thread
    t = time.perf_counter()
    O.calc_transforms(mode, scheme_id, X.nmem, md, 1.0, ob, S, s, lon_all[jj0, :8], lat_all[jj0, :8],
                      use_ref=trees[0] if use_ref else False)
    per = (time.perf_counter() - t) / 8
    # stride rows for about args.cpu_seconds of wall time on nth threads
    nrow_calc = int(max(2, min(X.nj_s - jj0, args.cpu_seconds * nth / max(per, 1e-6) / X.ni_s)))
    lon = lon_all[jj0:jj0 + nrow_calc]
    lat = lat_all[jj0:jj0 + nrow_calc]
    log("CPU baseline: %s, %d threads, %d stride rows (%d nodes, %.1f ms per node on one core)"
        % (kind, nth, nrow_calc, lon.size, 1e3 * per))
    t = time.perf_counter()
    ref = O.calc_transforms(mode, scheme_id, X.nmem, md, 1.0, ob, S, s, lon.ravel(), lat.ravel(),
                            use_ref=trees, threads=nth)
    dt = time.perf_counter() - t
    if use_ref:
        for tr in trees:
            tr.close()
    pbar = ref["stats"]["nlobs_sum"] / max(ref["stats"]["ncell"], 1)
    what = ("reference calc/calcs.c + common/kdtree.c (MKL sequential, %d threads; one set of kd-trees "
            "per thread, built outside the timing) driven by the restated das_calctransforms loop" % nth
            if use_ref else "oracle restatement (plain C, %d threads)" % nth)
    res = dict(value=lon.size / dt, unit="local analyses/s", cores=nth, kind=kind,
               sample="%d stride rows x %d nodes from the middle of the tile (mean %.0f local obs), %s, %.1f s"
                      % (nrow_calc, X.ni_s, pbar, what, dt))
    del S, s

    # the engine on the same band: FP64 parity and mixed error against the reference
    dotT = mode != enkfg.MODE_ENOI
    par = {}
    engines = {}
    for prec in (enkfg.PREC_FP64, enkfg.PREC_MIXED):
        E = enkfg.Engine(X.nmem, precision=prec, pertype_diag=True, keep_fp64=True, **engine_kw)
        E.set_obstypes(X.obstypes)
        E.set_obs(X.obs_type, X.obs_status, X.obs_lon, X.obs_lat, X.obs_estd)
        E.set_S(X.HA, X.innov, standardised=False)
        E.calc_transforms(lon, lat, stride=X.stride)
        T, w, T64, w64 = E.transforms(fp64=True)
        pn, pd, ps = E.diag()
        tag = "fp64" if prec == enkfg.PREC_FP64 else "mixed"
        d = dict(w64=relerr(w64, ref["w64"]), nlobs_equal=bool(np.array_equal(pn, ref["pnlobs"])))
        if dotT:
            d["T64"] = relerr(T64, ref["T64"])
            d["T_fp32_max_abs"] = float(np.abs(T.astype(np.float64) - ref["T"]).max())
        nz = ref["pdfs"] != 0
        d["dfs"] = float(np.abs(pd[nz] - ref["pdfs"][nz]).max() / np.abs(ref["pdfs"][nz]).max())
        par[tag] = d
        engines[prec] = E
        del T64, w64
    res["parity"] = dict(nodes=int(lon.size), against=kind, note="max |engine - reference| / max |reference| "
                         "over the band; FP64 bar 1e-10", **par)

    # update: model rows of the first nrow_upd stride rows of the band
    if dotT and not args.no_update_baseline:
        m = X.nmem
        r0 = jj0 * X.stride
        t = time.perf_counter()
        one = band_buffers_one(X, 0, r0, X.stride + 1)
        Tg = ref["T"].reshape(nrow_calc, X.ni_s, md, m)
        wg = ref["w"].reshape(nrow_calc, X.ni_s, m)
        nl1 = np.ascontiguousarray(X.nlevels[r0:r0 + X.stride + 1])
        t = time.perf_counter()
        O.update_fields(mode, m, md, X.stride, 0, 0, nl1, Tg[:2], wg[:2], X.fields[:1], [one], 0, X.stride)
        per_row = (time.perf_counter() - t) / (X.stride + 1)     # one field, one model row
        want_rows = args.cpu_seconds * nth / max(per_row * X.nfield, 1e-9)
        nrow_upd = int(max(1, min(nrow_calc - 1, want_rows // X.stride)))
        rows = nrow_upd * X.stride + 1
        log("CPU update baseline: %d model rows x %d fields (%.2f ms per field-row on one core)"
            % (rows, X.nfield, 1e3 * per_row))
        bufs = band_buffers(X, r0, rows)
        nl = np.ascontiguousarray(X.nlevels[r0:r0 + rows])
        Tu, wu = Tg[:nrow_upd + 1], wg[:nrow_upd + 1]
        cref = [b.copy() for b in bufs]
        errs = []

        def work(k):
            fs = list(range(k, X.nfield, nth))
            try:
                O.update_fields(mode, m, md, X.stride, 0, 0, nl, Tu, wu, [X.fields[f] for f in fs],
                                [cref[f] for f in fs], 0, rows - 1)
            except Exception as e:      # reported below, never silently
                errs.append(e)
        t = time.perf_counter()
        ths = [threading.Thread(target=work, args=(k,)) for k in range(nth)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        dtu = time.perf_counter() - t
        if errs:
            raise errs[0]
        wet = 0
        for f in X.fields:
            thr = 0 if f.get("issurfacevar") else f["level"]
            wet += int((nl > thr).sum())
        cm = X.nfield * rows * X.ni * md
        res_u = dict(value=cm / dtu, unit="cell·member updates/s", cores=nth, kind="port",
                     sample="%d model rows x %d columns x %d fields (%d wet cell-fields), restated FP32 "
                            "das_updatefields (oracle, plain C), fields split over %d threads, %.1f s"
                            % (rows, X.ni, X.nfield, wet, nth, dtu))
        # the engine's update of the same rows from host buffers (PCIe included), FP64 and mixed
        for prec in (enkfg.PREC_FP64, enkfg.PREC_MIXED):
            E = engines[prec]
            E.calc_transforms(lon[:nrow_upd + 1], lat[:nrow_upd + 1], stride=X.stride)
            gb = [b.copy() for b in bufs]
            E.sync()
            t = time.perf_counter()
            E.update_fields(nl, X.stride, 0, 0, X.fields, gb, 0, rows - 1)
            dtg = time.perf_counter() - t
            A = np.stack([b[:md] for b in gb]).astype(np.float64)          # [f][e][rows][ni]
            B = np.stack([b[:md] for b in cref]).astype(np.float64)
            F0 = np.stack([b[:md] for b in bufs]).astype(np.float64)
            ma, mb, mf = A.mean(1), B.mean(1), F0.mean(1)
            sa, sb = A.std(1), B.std(1)
            d = dict(host_buffer_update_ms=1e3 * dtg, host_buffer_cell_member_per_s=cm / dtg,
                     field_max_rel=float(np.abs(A - B).max() / max(np.abs(B).max(), 1e-30)),
                     mean_rel=float(np.abs(ma - mb).max() / np.abs(mb).max()),
                     mean_rel_to_increment=float(np.abs(ma - mb).max() / max(np.abs(mb - mf).max(), 1e-30)),
                     spread_rel=float(np.abs(sa - sb).max() / np.abs(sb).max()))
            res_u["fp64" if prec == enkfg.PREC_FP64 else "mixed"] = d
            del A, gb
        res_u["note"] = ("engine transforms (FP64 / mixed) and update (FP32 / FP16 hi-lo MFMA) of the band "
                         "against the reference transforms with the restated update; host_buffer_* times "
                         "enkfg_update_fields with host buffers, PCIe transfers included")
        res["update"] = res_u
    for E in engines.values():
        E.close()
    return res


def band_buffers_one(X, f, r0, rows):
    import numpy as np
    m, ni = X.nmem, X.ni
    a = X.E.read((f * X.ncell + r0 * ni) * m, rows * ni * m).reshape(rows * ni, m)
    return np.ascontiguousarray(a.T).reshape(m, rows, ni)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=["c2", "c4"], default="c2",
                    help="c2: the 1500x1800x51 tile, m=96 (FP64 value, mixed beside it); c4: hybrid 96+48 "
                         "members on 1200x1200x51 with clustered obs (mixed value)")
    ap.add_argument("--scheme", choices=["denkf", "etkf"], default="denkf")
    ap.add_argument("--size", choices=["full", "small"], default="full")
    ap.add_argument("--order", choices=["random", "prep"], default="random",
                    help="obs order: as generated, or sorted as enkf_prep writes super-obs (type, i, j, k)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-update-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="wall time of each CPU baseline sample")
    ap.add_argument("--precision", choices=["fp64", "both"], default="both",
                    help="both: the FP64 path and the mixed path")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run two plain device copies of GIB GiB (the byte counters' reference reading) and exit")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import synth_gpu
        nb = int(args.copy_probe * (1 << 30)) // 16 * 16
        dt = synth_gpu.copy_probe(nb)
        log("copy probe: %d bytes read and %d written in %.3f ms (%.1f GB/s each way)" % (nb, nb, 1e3 * dt,
                                                                                          nb / dt / 1e9))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        log("bench.py measures one GPU: the path would shard by stride-row bands balanced by local "
            "obs counts, but only the one-GPU run is measured.  Use --gpus 1.")
        sys.exit(2)

    import numpy as np
    import enkfg
    import synth_gpu

    c4 = args.config == "c4"
    if c4:
        tile = dict(nj=1200, ni=1200, nk=51, nmem=144, nmem_dynamic=96, median=400.0, pmax=20000.0)
        if args.size == "small":
            tile.update(nj=300, ni=300, nk=10, pmax=5000.0)
        mode = enkfg.MODE_HYBRID
    else:
        if args.size == "full":
            tile = dict(nj=1500, ni=1800, nk=51, nmem=96, n_sst=4_200_000, n_sla=2_100_000, n_prof=6_900)
        else:
            tile = dict(nj=300, ni=360, nk=10, nmem=96, n_sst=170_000, n_sla=84_000, n_prof=280)
        tile["order"] = args.order
        mode = enkfg.MODE_ENKF
    scheme_id = enkfg.SCHEME_DENKF if args.scheme == "denkf" else enkfg.SCHEME_ETKF
    log("generating the synthetic %s problem %s on the GPU ..." % (args.config, tile))
    t0 = time.perf_counter()
    X = synth_gpu.make_clustered_tile(**tile, seed=0) if c4 else synth_gpu.make_tile(**tile, seed=0)
    log("  done in %.1f s: %d obs, %d x %d nodes, %d fields" % (time.perf_counter() - t0, X.nobs, X.nj_s, X.ni_s,
                                                                 X.nfield))
    m = X.nmem
    md = X.nmem_dynamic if c4 else m
    engine_kw = dict(nmem_dynamic=md, mode=mode, scheme=scheme_id)
    upd_flags = enkfg.UPD_HYBRID_CONDITION if c4 else 0
    n_an = X.nj_s * X.ni_s
    nfield, ncell = X.nfield, X.ncell
    # a band of stride rows in the middle of the tile, for the mixed-versus-FP64 comparison
    jb = X.nj_s // 2
    sample = (X.node_lon[jb:jb + 8], X.node_lat[jb:jb + 8])
    pname = {enkfg.PREC_FP64: "fp64", enkfg.PREC_MIXED: "mixed"}

    spread_ms = None

    def run(precision):
        """warmup + timed steps of one precision; returns rates, phases and the
        transforms of the comparison band"""
        nonlocal spread_ms
        E = enkfg.Engine(m, precision=precision, pertype_diag=True, keep_fp64=False, **engine_kw)
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
                E.ens_attach(X.nlevels, X.stride, 0, 0, X.fields, X.E.ptr)
                if precision == enkfg.PREC_FP64 and spread_ms is None:
                    # das_writespread of the forecast: every field, before any update
                    dsp = synth_gpu.DeviceArray((nfield, X.nj, X.ni), np.float32)
                    for _ in range(2):
                        E.ens_spread(0, nfield, (X.nj, X.ni), dev_out=dsp.ptr)
                    spread_ms = E.spread_ms()
                    dsp.free()
                    log("  [fp64] forecast spread of %d fields: kernel %.2f ms" % (nfield, spread_ms))
            if w < args.warmup:
                E.ens_update(upd_flags)
                ph[3] = E.timings()[3]
                log("  [%s] warmup %d: phases (ms) loc %.1f gram %.1f factor %.1f update %.1f"
                    % (pname[precision], w, *ph))
        E.sync()
        phases = np.zeros(4)
        tcalc = tupd = 0.0
        t0 = time.perf_counter()
        for k in range(args.steps):
            a = time.perf_counter()
            E.calc_transforms(X.node_lon, X.node_lat, stride=X.stride)
            b = time.perf_counter()
            ph = E.timings()
            E.ens_update(upd_flags)
            c = time.perf_counter()
            ph[3] = E.timings()[3]
            phases += np.array(ph)
            tcalc += b - a
            tupd += c - b
            log("  [%s] step %d: calc %.1f ms update %.1f ms (loc %.1f gram %.1f factor %.1f update-kernel %.1f)"
                % (pname[precision], k, 1e3 * (b - a), 1e3 * (c - b), *ph))
        E.sync()
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
        E = enkfg.Engine(m, precision=precision, pertype_diag=True, keep_fp64=False, graph=True, **engine_kw)
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

    R = run(enkfg.PREC_FP64)
    Rm = run(enkfg.PREC_MIXED) if args.precision == "both" else None
    graph_ms = None if args.no_graph else run_graph(enkfg.PREC_MIXED if c4 else enkfg.PREC_FP64)
    V = Rm if (c4 and Rm is not None) else R     # the configuration's own precision
    K = args.steps
    phases = V["phases"]
    st = V["stats"]
    calc_rate, upd_rate = V["calc_rate"], V["upd_rate"]
    prec_tag = "mixed" if V is Rm else "fp64"
    if c4:
        tag = "c4_%s_%s" % (prec_tag, args.scheme) if args.size == "full" else "small_c4_%s" % args.scheme
    else:
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
    mt = (m + 15) // 16
    kt, ot = (m + 15) // 16, (md + 15) // 16
    full = (m == 16 * kt and md == 16 * ot)
    uname = "update_tile_kernel<%d, %d, %s, %s, %s>" % (kt, kt if not (kt == 9 and ot == 6) else ot,
                                                      "true" if V is Rm else "false",
                                                      "true" if m % 4 == 0 else "false", "true" if full else "false")
    uach = B_upd / (phases[3] * 1e-3)
    upd_roof = dict(bound="hbm", achieved=uach / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s", frac=uach / PEAK_HBM,
                    traffic=profiled_traffic(tag, uname), kernel=uname, launches_per_step=1,
                    avg_launch_ms=phases[3],
                    algorithmic="4*m*(2*wet cell-fields + dry cell-fields) + 4*N_an*(m*m_d+m) = %.3e B per launch"
                                % B_upd)
    if dom == "update":
        roof = upd_roof
    elif dom == "gram":
        nl = V["launches"]
        if V is Rm:
            F = 3 * F_gram
            ach = F / (phases[1] * 1e-3)
            kname = ("gram_mixed_item_kernel<%d, *>" % mt if m <= 96 else "gram_mixed_kernel<%d>" % mt)
            roof = dict(bound="mfma", achieved=ach / 1e12, peak=2500.0, unit="TFLOP/s", frac=ach / 2.5e15,
                        traffic=profiled_traffic(tag, kname), kernel=kname + " (BF16 MFMA, 3 hi/lo products)",
                        launches_per_step=nl, avg_launch_ms=phases[1] / max(nl, 1),
                        algorithmic="3 * sum_nodes p*(m(m+1)+2m) = %.3e flop per step over %d launches" % (F, nl))
        else:
            ach = F_gram / (phases[1] * 1e-3)
            # the item kernel in every tile shape the autotuner chose (per
            # obs-count bucket): the FP64 Gram phase is their launches together
            kname = "gram_item_kernel<%d, *>" % mt
            roof = dict(bound="mfma", achieved=ach / 1e12, peak=PEAK_FP64 / 1e12, unit="TFLOP/s",
                        frac=ach / PEAK_FP64, traffic=profiled_traffic(tag, kname), kernel=kname + " (FP64 MFMA)",
                        launches_per_step=nl, avg_launch_ms=phases[1] / max(nl, 1),
                        algorithmic="sum_nodes p*(m(m+1)+2m) = %.3e flop per step over %d launches"
                                    % (F_gram, nl))
    else:
        roof = dict(bound="latency", achieved=None, peak=None, unit=None, frac=None, traffic=None,
                    kernel=dom, algorithmic="dominant phase is %s" % dom)

    if c4:
        workload = ("hybrid %d+%d members, grid %dx%dx%d, %d clustered obs (lognormal local counts, median %g, "
                    "max %g, sigma %.2f), %s, %s, stride %d, locrad 150 km, %d local analyses, %d fields"
                    % (md, m - md, X.nk, X.nj, X.ni, X.nobs, tile["median"], tile["pmax"], X.sigma,
                       args.scheme.upper(), prec_tag.upper(), X.stride, n_an, nfield))
    else:
        workload = ("tile %dx%dx%d, m=%d, %d obs (%s order), %s, %s, stride %d, locrad 250 km, %d local analyses, "
                    "%d fields" % (X.nk, X.nj, X.ni, m, X.nobs, args.order, args.scheme.upper(), prec_tag.upper(),
                                   X.stride, n_an, nfield))
    res = dict(metric=METRIC, value=calc_rate, unit="local analyses/s", n_gpus=1, steps=K, warmup=args.warmup,
               ms_per_step=1e3 * (V["wall"]) / K, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64" if V is R else "bf16 hi/lo + f64",
               data="synthetic (GPU-generated Gaussian random fields; no files)",
               config=dict(workload=workload, grid=[X.nk, X.nj, X.ni], nmem=m, nmem_dynamic=md, nobs=int(X.nobs),
                           n_an=n_an, nfield=nfield, precision=prec_tag, scheme=args.scheme,
                           mean_local_obs=p_sum / n_an, max_local_obs=st["nlobs_max"]),
               update_cell_member_per_s=upd_rate,
               phases_ms=dict(localise=phases[0], gram=phases[1], factor=phases[2], update=phases[3],
                              calc_wall=1e3 * V["tcalc"], update_wall=1e3 * V["tupd"]),
               kd_tree_build_s=V["t_tree"],
               roofline=roof,
               update_roofline=upd_roof,
               calc_buckets=V["tuning"])
    if spread_ms:
        B_sp = 4.0 * nfield * ncell * (m + 1)
        res["spread"] = dict(kernel="spread_cells_kernel<true>", note="das_writespread of the forecast, all fields",
                             ms=spread_ms, achieved_GBps=B_sp / (spread_ms * 1e-3) / 1e9,
                             frac=B_sp / (spread_ms * 1e-3) / PEAK_HBM,
                             algorithmic="4*(m+1) B per cell-field = %.3e B per launch" % B_sp)
    if graph_ms is not None:
        res["graph"] = dict(note="the same calc pass with the Gram/factor batches replayed as one captured "
                                 "HIP graph (ENKFG_FLAG_GRAPH)", calc_ms=graph_ms,
                            local_analyses_per_s=n_an / (graph_ms * 1e-3),
                            speedup_vs_stream_launches=1e3 * V["tcalc"] / graph_ms)
    if Rm is not None:
        Tb, wb, pdb = R["band"]
        Tm, wm, pdm = Rm["band"]
        nz = pdb != 0
        F_mix = 3 * p_sum * (m * (m + 1) + 2 * m)       # three split products
        pm = Rm["phases"]
        res["mixed"] = dict(
            precision="S^T S as BF16 hi/lo MFMA (3 products, FP32 accumulation); factorisation FP64; "
                      "update X^f T as FP16 hi/lo MFMA (3 products, FP32 accumulation)",
            local_analyses_per_s=Rm["calc_rate"], update_cell_member_per_s=Rm["upd_rate"],
            calc_speedup_vs_fp64=Rm["calc_rate"] / R["calc_rate"],
            step_speedup_vs_fp64=(R["tcalc"] + R["tupd"]) / (Rm["tcalc"] + Rm["tupd"]),
            phases_ms=dict(localise=pm[0], gram=pm[1], factor=pm[2], update=pm[3],
                           calc_wall=1e3 * Rm["tcalc"], update_wall=1e3 * Rm["tupd"]),
            gram_bf16_tflops=F_mix / (pm[1] * 1e-3) / 1e12,
            calc_buckets=Rm["tuning"],
            error_vs_fp64=dict(
                sample="8 stride rows (%d nodes) in the middle of the tile" % (8 * X.ni_s),
                T=relerr(Tm, Tb), w=relerr(wm, wb),
                dfs=float(np.abs(pdm[nz] - pdb[nz]).max() / np.abs(pdb[nz]).max())))
        res["fp64"] = dict(local_analyses_per_s=R["calc_rate"], update_cell_member_per_s=R["upd_rate"],
                           phases_ms=dict(localise=R["phases"][0], gram=R["phases"][1], factor=R["phases"][2],
                                          update=R["phases"][3], calc_wall=1e3 * R["tcalc"],
                                          update_wall=1e3 * R["tupd"]))
    if not args.no_cpu_baseline:
        log("CPU baseline and band parity ...")
        cb = cpu_side(X, args, mode, scheme_id, md, engine_kw)
        res["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        res["cpu_baseline"]["parity"] = cb["parity"]
        if "update" in cb:
            res["cpu_baseline_update"] = cb["update"]
        log("  %s" % json.dumps(cb))
    X.E.free()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
