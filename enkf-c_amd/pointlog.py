"""Point logs of the calc: das_calcpointlogtransforms (calc/transforms.c:1640-1770),
with the files of das_createplog and das_writeplogtransform
(common/pointlog.c:83-346), named as das_getfname_plog does
(common/dasystem.c:585-588, prefix "pointlog" from common/global.h:36).

The reference computes the transforms of the log points one by one on the
CPU.  Here all log points of a grid are one row of nodes handed to the
engine's calc: the same localisation, Gram, factorisation and relaxation to
prior as the grid's own nodes run for them on the GPU, in one call.  The
files are then written on the host with enkfio.

A calc replaces the engine's transforms, so the point logs run after the
grid's transforms have been fetched or written; enkf_calc does the same
(calc/enkf_calc.c:396-405: das_calctransforms, then the point logs).  The
engine must keep the FP64 transforms (keep_fp64=True): the files hold T and
w as doubles.

One difference from the reference: for a point without local obs the
reference defines w and T but writes no values (pointlog.c:413-420 writes
only when ploc > 0), so a reader sees fill values; here they hold what the
calc gives such a point, w = 0 and T = I.
"""
import os

import numpy as np

import enkfg
import enkfio


def plog_fname(lon, lat):
    """das_getfname_plog (common/dasystem.c:585-588)"""
    return "pointlog_%.3f,%.3f.nc" % (lon, lat)


def _gridstr(gid, ngrid):
    """get_gridstr (common/pointlog.c:71-77)"""
    return "" if ngrid == 1 else "-%d" % gid


def write_pointlog(path, mode, nmem, nmem_dynamic, plon, plat, ids, lc, obs, grids, transforms,
                   scheme=enkfg.SCHEME_DENKF, alpha=1.0, gamma=1.0, obstypes=None, instruments=None,
                   datestr=None):
    """One point log file: das_createplog (common/pointlog.c:83-346) followed by
    das_writeplogtransform (348-426) for every grid.

    ids, lc        the point's local obs over all domains and their taper
                   coefficients (obs_findlocal with no domain)
    obs            dict of per-observation arrays (enkfio.read_obs keys: lon,
                   lat, depth, value, estd, fi, fj, fk, type, instrument, time)
    grids          list of dicts: name, and optionally domain, fi, fj, nk,
                   model_depth, alias (alias grids get no transforms)
    transforms     per grid, None (alias) or a dict with s [p], S [m][p]
                   (the local standardised innovations and anomalies, times
                   the taper), w [m] and T [m_d][m] (None in EnOI mode)
    obstypes       list of dicts with name and optionally rfactor, locrad,
                   locweight, gridid, domains (the type attributes)
    instruments    list of instrument names (ids are their positions)
    """
    ngrid = len(grids)
    ploc = len(ids)
    with enkfio.NcWriter(path) as f:
        if mode == enkfg.MODE_HYBRID:
            f.dim("m1", nmem_dynamic)
            f.dim("m2", nmem)
            mrow, mcol = "m1", "m2"
        else:
            f.dim("m", nmem)
            mrow, mcol = "m", "m"
        f.dim("p", ploc)
        f.var("dimension_reference", np.int8, [])
        vnames = [("obs_ids", np.int32), ("obs_lcoeffs", np.float32), ("obs_lon", np.float32),
                  ("obs_lat", np.float32), ("obs_depth", np.float32), ("obs_val", np.float32),
                  ("obs_estd", np.float32), ("obs_fi", np.float32), ("obs_fj", np.float32),
                  ("obs_fk", np.float32), ("obs_type", np.int32), ("obs_inst", np.int32),
                  ("obs_time", np.float32)]
        for name, dt in vnames:
            f.var(name, dt, ["p"])
        f.attr("long_name", "observation time", "obs_time")
        if datestr is not None:
            f.attr("units", "days from %s" % datestr, "obs_time")
        for otid, ot in enumerate(obstypes or []):
            if ot.get("statsonly", 0):
                continue
            f.attr(ot["name"], otid, "obs_type")
        for ot in obstypes or []:
            if ot.get("statsonly", 0):
                continue
            nm = ot["name"]
            f.attr("RFACTOR_" + nm, float(ot.get("rfactor", 1.0)), "obs_type")
            lr = list(ot.get("locrad", []))
            if lr:
                f.attr("LOCRAD_" + nm, np.asarray(lr, np.float64), "obs_type")
                lw = list(ot.get("locweight", [1.0] * len(lr)))
                f.attr("WEIGHT_" + nm, np.asarray(lw, np.float64) / sum(lw), "obs_type")
            f.attr("GRIDID_" + nm, int(ot.get("gridid", 0)), "obs_type")
            if ot.get("domains"):
                f.attr("DOMAINS_" + nm, " ".join(ot["domains"]), "obs_type")
        for gid, g in enumerate(grids):
            gv = "grid" + _gridstr(gid, ngrid)
            f.var(gv, np.int32, [])
            f.attr("id", gid, gv)
            f.attr("name", g["name"], gv)
            f.attr("domain", g.get("domain", "Default"), gv)
            if "fi" in g:
                f.attr("fi", float(g["fi"]), gv)
                f.attr("fj", float(g["fj"]), gv)
            if "nk" in g:
                f.attr("nk", int(g["nk"]), gv)
            if "model_depth" in g:
                f.attr("model_depth", np.float32(g["model_depth"]), gv)
        for iid, name in enumerate(instruments or []):
            f.attr(name, iid, "obs_inst")
        f.attr("lon", float(plon))
        f.attr("lat", float(plat))
        if mode in (enkfg.MODE_ENKF, enkfg.MODE_HYBRID):
            f.attr("MODE", "EnKF" if mode == enkfg.MODE_ENKF else "Hybrid")
            f.attr("SCHEME", "DEnKF" if scheme == enkfg.SCHEME_DENKF else "ETKF")
            f.attr("ALPHA", float(alpha))
        else:
            f.attr("MODE", "EnOI")
        if mode == enkfg.MODE_HYBRID:
            f.attr("GAMMA", float(gamma))
        f.attr("ngrids", ngrid)

        # das_writeplogtransform: per grid s, S (when the point has local
        # obs), w, and T (EnKF and hybrid)
        for gid, (g, tr) in enumerate(zip(grids, transforms)):
            if tr is None:
                continue
            gs = _gridstr(gid, ngrid)
            pl = len(tr["s"])
            if pl > 0:
                pd = "p" + gs
                if pd != "p":
                    f.dim(pd, pl)
                    f.attr(pd, "number of observations on grid %d" % gid, "dimension_reference")
                f.var("s" + gs, np.float32, [pd])
                f.var("S" + gs, np.float32, [mcol, pd])
                # the reference sets both long names on s, the second replacing
                # the first (pointlog.c:391-395)
                f.attr("long_name", "standardised observation anomalies", "s" + gs)
            f.var("w" + gs, np.float64, [mcol])
            f.attr("long_name", 'increment coefficients on grid %d ("%s")' % (gid, g["name"]), "w" + gs)
            if tr.get("T") is not None:
                f.var("T" + gs, np.float64, [mrow, mcol])
        f.enddef()

        if ploc > 0:
            ids = np.asarray(ids)
            f.put("obs_ids", ids.astype(np.int32))
            f.put("obs_lcoeffs", np.asarray(lc, np.float64).astype(np.float32))
            for name, key in (("obs_lon", "lon"), ("obs_lat", "lat"), ("obs_depth", "depth"),
                              ("obs_val", "value"), ("obs_estd", "estd"), ("obs_fi", "fi"),
                              ("obs_fj", "fj"), ("obs_fk", "fk"), ("obs_type", "type"),
                              ("obs_inst", "instrument"), ("obs_time", "time")):
                v = obs.get(key)
                dt = np.int32 if name in ("obs_type", "obs_inst") else np.float32
                f.put(name, np.zeros(ploc, dt) if v is None else np.asarray(v)[ids].astype(dt))
        for gid, tr in enumerate(transforms):
            if tr is None:
                continue
            gs = _gridstr(gid, ngrid)
            if len(tr["s"]) > 0:
                f.put("s" + gs, np.asarray(tr["s"], np.float64).astype(np.float32))
                f.put("S" + gs, np.asarray(tr["S"], np.float64).astype(np.float32))
            f.put("w" + gs, np.asarray(tr["w"], np.float64))
            if tr.get("T") is not None:
                f.put("T" + gs, np.asarray(tr["T"], np.float64))


def calc_pointlogs(eng, points, obs, grids=None, geographic=1, outdir=".", obstypes=None,
                   instruments=None, datestr=None, S=None, s=None):
    """das_calcpointlogtransforms (calc/transforms.c:1640-1770) for the log
    points [(lon, lat), ...]: the transforms of every point on the GPU (one
    calc per grid over all points), then one file per point.  eng holds the
    obs and S (set_obs, set_S); grids as in write_pointlog, each with an
    optional `visible` array (the obs types visible from the grid's domain).
    S, s: the standardised S [m][nobs] and s, when the caller already holds
    them (else read back from the engine).  Returns the file paths."""
    if not getattr(eng, "keep_fp64", False):
        raise ValueError("point logs need an engine created with keep_fp64=True")
    grids = grids or [dict(name="grid")]
    pts = np.asarray(points, np.float64).reshape(-1, 2)
    npt = len(pts)
    lon = pts[:, 0].reshape(1, npt).copy()
    lat = pts[:, 1].reshape(1, npt).copy()
    if S is None or s is None:
        S, s = eng.get_S()
    enoi = eng.mode == enkfg.MODE_ENOI

    # the local obs of every point over all domains (das_createplog)
    eng.localise(lon, lat, geographic=geographic)
    cnt = eng.local_counts()
    lists = [eng.localobs(k, cnt[k]) for k in range(npt)]
    lists = [(ids[:n].copy(), lc[:n].copy()) for (ids, lc), n in zip(lists, cnt)]

    per_grid = []
    for g in grids:
        if g.get("alias", False):
            per_grid.append(None)
            continue
        eng.calc_transforms(lon, lat, stride=1, geographic=geographic, visible=g.get("visible"))
        gc = eng.local_counts()
        _, _, T64, w64 = eng.transforms(fp64=True)
        trs = []
        for k in range(npt):
            ids, lc = eng.localobs(k, gc[k])
            ids, lc = ids[:gc[k]], lc[:gc[k]]
            # Sloc[e][o] = S[e][lobs[o]] * lcoeffs[o], sloc[o] = s[lobs[o]] * lcoeffs[o]
            # in double (transforms.c:1718-1727)
            trs.append(dict(s=s[ids] * lc, S=S[:, ids].astype(np.float64) * lc[None, :], w=w64[k],
                            T=None if enoi else T64[k]))
        per_grid.append(trs)

    paths = []
    for k in range(npt):
        path = os.path.join(outdir, plog_fname(pts[k, 0], pts[k, 1]))
        write_pointlog(path, eng.mode, eng.nmem, eng.md, pts[k, 0], pts[k, 1], lists[k][0], lists[k][1], obs,
                       grids, [None if pg is None else pg[k] for pg in per_grid], scheme=eng.scheme,
                       alpha=eng.alpha, gamma=eng.gamma, obstypes=obstypes, instruments=instruments,
                       datestr=datestr)
        paths.append(path)
    return paths
