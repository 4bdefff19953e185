"""The adjoint of the analysis mean (enkfg_adjoint) against an FP64 reference
gradient and central finite differences.

J = sum over a box of cells and the fields of the ensemble of
(x^a - x^t)^2, with x^a = x^f + A^f w the analysis mean of update.c:381-410
and w interpolated from the stride-grid transforms.  The GPU forward runs
through the C ABI (FP64 calc path); its backward is enkfg_adjoint, wrapped in
a torch.autograd.Function here only.  The reference is the same forward
written in float64 torch with no FP32 rounding of S, differentiated by torch
autograd, and central differences of that forward (h = 1e-6, relative for
the variances).  Bounds: max |gpu - reference| <= 1e-4 max |reference| for
dJ/dy and dJ/dR; dJ/dinflation is exactly 0 (inflation scales anomalies
only).
"""
import math

import numpy as np
import pytest
import torch

import enkfg
import oracle as O
import synth

pytestmark = pytest.mark.gpu


class Case:
    def __init__(self, seed=31):
        P = synth.make_problem(seed=seed, nj=24, ni=30, nk=2, nmem=24, nobs=700, locrad=80.0, stride=3, profiles=10)
        self.P = P
        m = P.nmem
        self.m = m
        self.nobs = P.nobs
        self.rf = np.array([P.obstypes[t]["rfactor"] for t in P.obs_type], np.float64)
        ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
        self.lists = [ob.findlocal(lo, la) for lo, la in zip(P.node_lon.ravel(), P.node_lat.ravel())]
        F = [P.ens["temp"][:, k] for k in range(P.nk)] + [P.ens["salt"][:, k] for k in range(P.nk)]
        self.fields = [dict(level=k) for k in range(P.nk)] * 2
        Ef = np.stack(F).astype(np.float32)                        # [nfield][m][nj][ni]
        self.Ecell = np.ascontiguousarray(Ef.reshape(len(F), m, -1).transpose(0, 2, 1))   # [f][cell][m]
        X = self.Ecell.astype(np.float64)
        self.xf = X.mean(2)
        self.Af = torch.tensor(X - self.xf[:, :, None])
        self.xf = torch.tensor(self.xf)
        xt = [P.truth["temp"][k] for k in range(P.nk)] + [P.truth["salt"][k] for k in range(P.nk)]
        self.xt = torch.tensor(np.stack(xt).reshape(len(xt), -1).astype(np.float64))
        box = np.zeros((P.nj, P.ni))
        box[6:17, 9:22] = 1.0
        self.box = torch.tensor(np.broadcast_to(box.ravel(), self.xt.shape).copy())
        self.B = torch.tensor(self.interp_matrix())
        self.HA = torch.tensor(P.HA.astype(np.float64))           # [m][nobs]

    def interp_matrix(self):
        """cells x nodes: the bilinear weights of update.c:195-272 (FP32
        weights, as real numbers)"""
        P = self.P
        s = P.stride
        B = np.zeros((P.nj * P.ni, P.nj_s * P.ni_s))
        for j in range(P.nj):
            for i in range(P.ni):
                jj, sj, ii, si = j // s, j % s, i // s, i % s
                jj2, ii2 = min(jj + 1, P.nj_s - 1), min(ii + 1, P.ni_s - 1)
                wj2 = np.float32(sj) / np.float32(s)
                wi2 = np.float32(si) / np.float32(s)
                bj = (1.0, 0.0) if sj == 0 else (float(np.float32(1) - wj2), float(wj2))
                bi = (1.0, 0.0) if si == 0 else (float(np.float32(1) - wi2), float(wi2))
                c = j * P.ni + i
                for (nj_, bjv) in ((jj, bj[0]), (jj2, bj[1])):
                    for (ni_, biv) in ((ii, bi[0]), (ii2, bi[1])):
                        B[c, nj_ * P.ni_s + ni_] += bjv * biv
        return B

    def mean_from_w(self, W):
        """x^a = x^f + A^f w_cell, W: [node][m] float64 tensor"""
        Wc = self.B @ W                                          # [cell][m]
        return self.xf + (self.Af * Wc[None]).sum(2)

    def J(self, xa):
        return (((xa - self.xt) ** 2) * self.box).sum()

    def forward64(self, innov, R):
        """the calc in float64 torch, S not rounded to FP32"""
        m = self.m
        rf = torch.tensor(self.rf)
        W = []
        for ids, lc in self.lists:
            if len(ids) == 0:
                W.append(torch.zeros(m, dtype=torch.float64))
                continue
            ids_t = torch.tensor(ids, dtype=torch.long)
            lc_t = torch.tensor(lc)
            d = torch.sqrt(R[ids_t] * rf[ids_t]) * math.sqrt(m - 1)
            S = self.HA[:, ids_t].T / d[:, None] * lc_t[:, None]
            s = innov[ids_t] / d * lc_t
            M = torch.eye(m, dtype=torch.float64) + S.T @ S
            W.append(torch.linalg.solve(M, S.T @ s))
        return self.mean_from_w(torch.stack(W))


class GpuMean(torch.autograd.Function):
    """x^a from the GPU calc (forward) and enkfg_adjoint (backward)"""

    @staticmethod
    def forward(ctx, innov, R, case, eng, dev_E):
        P = case.P
        estd = np.sqrt(R.detach().numpy()).astype(np.float32)
        eng.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, estd)
        eng.set_S(P.HA, innov.detach().numpy(), standardised=False)
        eng.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
        _, _, _, w64 = eng.transforms(fp64=True)
        eng.ens_attach(P.nlevels, P.stride, 0, 0, case.fields, dev_E.data_ptr())
        ctx.case, ctx.eng = case, eng
        return case.mean_from_w(torch.tensor(w64.reshape(-1, case.m)))

    @staticmethod
    def backward(ctx, grad):
        dJ = grad.to(torch.float32).contiguous().cuda()
        gy, gR, gi = ctx.eng.adjoint(dJ.data_ptr(), ctx.case.nobs)
        torch.cuda.synchronize()
        ctx.ginfl = gi
        assert gi == 0.0
        return torch.tensor(gy), torch.tensor(gR), None, None, None


@pytest.fixture(scope="module")
def case():
    return Case()


def test_adjoint_against_autograd_and_finite_differences(case):
    print("case adjoint m=%d nobs=%d" % (case.m, case.nobs), flush=True)
    P = case.P
    R0 = torch.tensor(P.obs_estd.astype(np.float64) ** 2)
    innov0 = torch.tensor(P.innov)
    # reference gradient: float64 forward, torch autograd
    iv = innov0.clone().requires_grad_(True)
    Rv = R0.clone().requires_grad_(True)
    Jr = case.J(case.forward64(iv, Rv))
    Jr.backward()
    gy_ref, gR_ref = iv.grad.clone(), Rv.grad.clone()
    # GPU forward + enkfg_adjoint
    eng = enkfg.Engine(case.m, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF, keep_fp64=True)
    eng.set_obstypes(P.obstypes)
    dev_E = torch.tensor(case.Ecell, device="cuda")
    ig = innov0.clone().requires_grad_(True)
    Rg = R0.clone().requires_grad_(True)
    Jg = case.J(GpuMean.apply(ig, Rg, case, eng, dev_E))
    Jg.backward()
    eng.close()
    gy, gR = ig.grad, Rg.grad
    assert abs(Jg.item() - Jr.item()) <= 1e-5 * abs(Jr.item())
    ey = float((gy - gy_ref).abs().max() / gy_ref.abs().max())
    eR = float((gR - gR_ref).abs().max() / gR_ref.abs().max())
    print("  J %.6g  |dJ/dy| %.3g  |dJ/dR| %.3g  err vs autograd: y %.2g  R %.2g"
          % (Jr.item(), gy_ref.abs().max(), gR_ref.abs().max(), ey, eR), flush=True)
    assert (gy_ref != 0).sum() > 50, "the box must be sensitive to many observations"
    assert ey <= 1e-4 and eR <= 1e-4
    # central differences of the float64 forward on the most sensitive obs
    with torch.no_grad():
        for name, base, g in (("y", innov0, gy), ("R", R0, gR)):
            worst = 0.0
            for o in torch.argsort(g.abs(), descending=True)[:6].tolist():
                h = 1e-6 if name == "y" else 1e-6 * float(base[o])
                xp, xm = base.clone(), base.clone()
                xp[o] += h
                xm[o] -= h
                if name == "y":
                    fd = (case.J(case.forward64(xp, R0)) - case.J(case.forward64(xm, R0))) / (2 * h)
                else:
                    fd = (case.J(case.forward64(innov0, xp)) - case.J(case.forward64(innov0, xm))) / (2 * h)
                worst = max(worst, abs(float(fd) - float(g[o])) / float(g.abs().max()))
            print("  finite differences %s: max error %.2g" % (name, worst), flush=True)
            assert worst <= 1e-4


def test_adjoint_requires_the_calc_and_the_ensemble(case):
    eng = enkfg.Engine(case.m, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF)
    eng.set_obstypes(case.P.obstypes)
    dJ = torch.zeros(4, 10, device="cuda")
    with pytest.raises(enkfg.EnkfgError) as ei:
        eng.adjoint(dJ.data_ptr(), case.nobs)
    assert ei.value.code == enkfg.ERR_STATE
    eng.close()


def test_adjoint_c5_m96_against_autograd_and_finite_differences():
    """The adjoint at the C5 size: 32 layers x 600 x 600, m = 96, 1M surface
    obs, LOCRAD 200 km, stride 3 (40,000 local analyses).  J is the squared
    deviation of the analysis mean from a target over a 100 x 100 box of the
    top 10 layers of temp and salt (20 fields).  The GPU forward is the whole
    FP64 calc; the reference is the same forward in float64 torch over the
    1,225 stride nodes whose transforms reach the box, differentiated by
    autograd, and central differences (h = 1e-6, relative for the variances)
    of that forward on the most sensitive observations."""
    import time
    import synth_gpu
    print("case adjoint C5: 32x600x600, m=96, 1M obs", flush=True)
    X = synth_gpu.make_tile(nj=600, ni=600, nk=32, nmem=96, n_sst=1_000_000, n_sla=0, n_prof=0, corr_km=150.0,
                            locrad=200.0, obs_std=(0.3, 0.04, 0.05), seed=5)
    m, s_ = X.nmem, X.stride
    nobs, ncell = X.nobs, X.ncell
    eng = enkfg.Engine(m, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF, keep_fp64=True)
    eng.set_obstypes(X.obstypes)
    eng.set_obs(X.obs_type, X.obs_status, X.obs_lon, X.obs_lat, X.obs_estd)
    eng.set_S(X.HA, X.innov, standardised=False)
    eng.calc_transforms(X.node_lon, X.node_lat, stride=s_)
    _, _, _, w64 = eng.transforms(fp64=True)
    eng.ens_attach(X.nlevels, s_, 0, 0, X.fields, X.E.ptr)
    # the box and the nodes whose interpolation weights reach it
    j0, i0, nb = 250, 250, 100
    fsel = list(range(10)) + list(range(X.nk, X.nk + 10))
    rows = np.arange(j0, j0 + nb)
    cols = np.arange(i0, i0 + nb)
    cells = (rows[:, None] * X.ni + cols[None, :]).ravel()
    wet = np.stack([X.nlevels.ravel()[cells] > X.fields[f]["level"] for f in fsel])   # [f][cell]
    Af = np.stack([np.concatenate([X.E.read((f * ncell + j * X.ni + i0) * m, nb * m).reshape(nb, m) for j in rows])
                   for f in fsel]).astype(np.float64)                         # [f][cell][member]
    xf = Af.mean(2)
    Af -= xf[:, :, None]
    jjs = sorted(set(rows // s_) | set(np.minimum(rows // s_ + 1, X.nj_s - 1)))
    iis = sorted(set(cols // s_) | set(np.minimum(cols // s_ + 1, X.ni_s - 1)))
    nodes = np.array([jj * X.ni_s + ii for jj in jjs for ii in iis])
    pos = {int(n): k for k, n in enumerate(nodes)}
    Bm = np.zeros((len(cells), len(nodes)))
    one = np.float32(1.0)
    for c, (j, i) in enumerate((j, i) for j in rows for i in cols):
        jj, sj, ii, si = j // s_, j % s_, i // s_, i % s_
        jj2, ii2 = min(jj + 1, X.nj_s - 1), min(ii + 1, X.ni_s - 1)
        wj2 = np.float32(sj) / np.float32(s_)
        wi2 = np.float32(si) / np.float32(s_)
        bj = (1.0, 0.0) if sj == 0 else (float(one - wj2), float(wj2))
        bi = (1.0, 0.0) if si == 0 else (float(one - wi2), float(wi2))
        for nj_, bjv in ((jj, bj[0]), (jj2, bj[1])):
            for ni_, biv in ((ii, bi[0]), (ii2, bi[1])):
                Bm[c, pos[nj_ * X.ni_s + ni_]] += bjv * biv
    Bt = torch.tensor(Bm)
    Aft = torch.tensor(Af)
    xft = torch.tensor(xf)
    wett = torch.tensor(wet.astype(np.float64))
    jg, ig = np.meshgrid(rows, cols, indexing="ij")
    patt = (np.cos(2 * np.pi * ig / 50.0) * np.sin(2 * np.pi * jg / 70.0)).ravel()
    xt = torch.tensor(xf + np.array([0.3 if f < X.nk else 0.03 for f in fsel])[:, None] * patt[None, :])

    def J_of_W(W):
        xa = xft + (Aft * (Bt @ W)[None]).sum(2)
        return (((xa - xt) ** 2) * wett).sum(), xa

    # GPU: forward mean from w64, backward via enkfg_adjoint
    J_gpu, xa_gpu = J_of_W(torch.tensor(w64[nodes]))
    dJ = np.zeros((X.nfield, ncell), np.float32)
    g = (2.0 * (xa_gpu - xt) * wett).numpy().astype(np.float32)
    for k, f in enumerate(fsel):
        dJ[f, cells] = g[k]
    ddJ = synth_gpu.DeviceArray.from_host(dJ)
    t = time.perf_counter()
    gy, gR, gi = eng.adjoint(ddJ.ptr, nobs)
    t_adj = time.perf_counter() - t
    eng.close()
    assert gi == 0.0
    # reference: float64 forward over the listed nodes, autograd
    ob = O.OracleObs(X.obstypes, X.obs_type, X.obs_lon, X.obs_lat, X.obs_estd)
    lists = [ob.findlocal(X.node_lon.ravel()[n], X.node_lat.ravel()[n]) for n in nodes]
    HA = torch.tensor(X.HA.astype(np.float64))
    sq = math.sqrt(m - 1)

    def node_w(k, innov, R):
        ids, lc = lists[k]
        if len(ids) == 0:
            return torch.zeros(m, dtype=torch.float64)
        it = torch.tensor(ids, dtype=torch.long)
        d = torch.sqrt(R[it]) * sq
        lct = torch.tensor(lc)
        S = HA[:, it].T / d[:, None] * lct[:, None]
        sv = innov[it] / d * lct
        M = torch.eye(m, dtype=torch.float64) + S.T @ S
        return torch.linalg.solve(M, S.T @ sv)

    R0 = torch.tensor(X.obs_estd.astype(np.float64) ** 2)
    inn0 = torch.tensor(X.innov)
    iv = inn0.clone().requires_grad_(True)
    Rv = R0.clone().requires_grad_(True)
    t = time.perf_counter()
    W = torch.stack([node_w(k, iv, Rv) for k in range(len(nodes))])
    Jr, _ = J_of_W(W)
    Jr.backward()
    t_ref = time.perf_counter() - t
    gy_ref, gR_ref = iv.grad.numpy(), Rv.grad.numpy()
    ey = float(np.abs(gy - gy_ref).max() / np.abs(gy_ref).max())
    eR = float(np.abs(gR - gR_ref).max() / np.abs(gR_ref).max())
    print("  %d nodes reach the box, %d obs with dJ/dy != 0; J %.6g (GPU forward %.6g)" % (
        len(nodes), int((gy_ref != 0).sum()), Jr.item(), J_gpu.item()), flush=True)
    print("  enkfg_adjoint %.2f s (torch float64 reference %.1f s); error vs autograd: y %.2g  R %.2g"
          % (t_adj, t_ref, ey, eR), flush=True)
    assert abs(J_gpu.item() - Jr.item()) <= 1e-5 * abs(Jr.item())
    assert (gy_ref != 0).sum() > 10000
    assert ey <= 1e-4 and eR <= 1e-4
    # central differences on the most sensitive obs: only the nodes whose
    # list holds the observation change
    with torch.no_grad():
        W0 = W.detach().clone()
        holds = {}
        for name, base, gg in (("y", inn0, gy), ("R", R0, gR)):
            worst = 0.0
            for o in np.argsort(-np.abs(gg))[:4]:
                o = int(o)
                ks = [k for k in range(len(nodes)) if o in holds.setdefault(k, set(lists[k][0].tolist()))]
                h = 1e-6 if name == "y" else 1e-6 * float(base[o])
                Js = []
                for sgn in (1.0, -1.0):
                    x = base.clone()
                    x[o] += sgn * h
                    Wp = W0.clone()
                    for k in ks:
                        Wp[k] = node_w(k, x, R0) if name == "y" else node_w(k, inn0, x)
                    Js.append(J_of_W(Wp)[0].item())
                fd = (Js[0] - Js[1]) / (2 * h)
                worst = max(worst, abs(fd - float(gg[o])) / float(np.abs(gg).max()))
            print("  finite differences %s: max error %.2g" % (name, worst), flush=True)
            assert worst <= 1e-4


@pytest.mark.parametrize("mode,m,md,alpha", [(enkfg.MODE_ENKF, 24, 24, 0.8), (enkfg.MODE_HYBRID, 30, 20, 1.0)])
def test_adjoint_ens_against_autograd(mode, m, md, alpha):
    """enkfg_adjoint_ens (DEnKF): J = sum over a box of cells, of 4 fields and
    of the dynamic members of (X^a - x^t)^2, so that J depends on the
    analysis anomalies as well as the mean.  The reference is the forward in
    float64 torch -- M, w = M^-1 S^T s and T = alpha/2 (I + M^-1) + (1 - alpha)
    I per node, their bilinear interpolation, v_a = A^f T^T (/ k_d when
    hybrid), the capped (ratio 1.0, 0.5) and plain (NaN ratio) inflation and
    X^a = infl v_a + x^f + A^f w -- differentiated by autograd with respect
    to the innovations, the observation-error variances and each field's
    inflation.  Bound: max |gpu - reference| <= 1e-4 max |reference|."""
    print("case adjoint_ens mode=%d m=%d md=%d alpha=%.1f" % (mode, m, md, alpha), flush=True)
    P = synth.make_problem(seed=41 + m, nj=24, ni=30, nk=2, nmem=m, nobs=700, locrad=80.0, stride=3, profiles=10)
    nobs = P.nobs
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    lists = [ob.findlocal(lo, la) for lo, la in zip(P.node_lon.ravel(), P.node_lat.ravel())]
    F = [P.ens["temp"][:, k] for k in range(P.nk)] + [P.ens["salt"][:, k] for k in range(P.nk)]
    nf = len(F)
    infl0 = [1.08, 1.05, 1.10, 1.30]
    ratio = [1.0, float("nan"), 0.5, 1.0]
    fields = [dict(level=k % P.nk, inflation=infl0[k], inf_ratio=ratio[k]) for k in range(nf)]
    Ecell = np.ascontiguousarray(np.stack(F).astype(np.float32).reshape(nf, m, -1).transpose(0, 2, 1))
    ncell = Ecell.shape[1]
    # forecast mean and anomalies as the update forms them (FP64 mean of the
    # dynamic members, FP32 anomalies)
    xf = Ecell[:, :, :md].astype(np.float64).mean(2)
    Af = (Ecell.astype(np.float64) - xf[:, :, None]).astype(np.float32).astype(np.float64)
    wet = np.stack([(P.nlevels.ravel() > fl["level"]) for fl in fields]).astype(np.float64)   # [f][cell]
    case = Case.__new__(Case)
    case.P = P
    Bm = torch.tensor(case.interp_matrix())                                  # [cell][node]
    box = np.zeros((P.nj, P.ni))
    box[5:18, 8:23] = 1.0
    boxt = torch.tensor(box.ravel()[None, :, None] * wet[:, :, None])      # [f][cell][1]
    xt = torch.tensor(np.stack([P.truth["temp"][k] for k in range(P.nk)] + [P.truth["salt"][k] for k in range(P.nk)])
                      .reshape(nf, -1)[:, :, None].astype(np.float64))
    HA = torch.tensor(P.HA.astype(np.float64))
    rf = torch.tensor(np.array([P.obstypes[t]["rfactor"] for t in P.obs_type], np.float64))
    hybrid = mode == enkfg.MODE_HYBRID
    k_d = math.sqrt((m - 1) / (md - 1)) if (hybrid and md > 1) else 1.0
    Aft = torch.tensor(Af)
    xft = torch.tensor(xf)

    def forward(innov, R, infl):
        Ws, Ts = [], []
        eye = torch.eye(m, dtype=torch.float64)
        for ids, lc in lists:
            if len(ids) == 0:
                Ws.append(torch.zeros(m, dtype=torch.float64))
                Ts.append(eye[:md])
                continue
            it = torch.tensor(ids, dtype=torch.long)
            d = torch.sqrt(R[it] * rf[it]) * math.sqrt(m - 1)
            lct = torch.tensor(lc)
            S = HA[:, it].T / d[:, None] * lct[:, None]
            sv = innov[it] / d * lct
            M = eye + S.T @ S
            Mi = torch.linalg.inv(M)
            Ws.append(Mi @ (S.T @ sv))
            Tn = 0.5 * (eye + Mi)
            Ts.append((alpha * (Tn - eye) + eye)[:md])                     # T[e1][e2], e1 < md
        W = torch.stack(Ws)
        T = torch.stack(Ts)                                                 # [node][md][m]
        Wc = Bm @ W                                                         # [cell][m]
        Tc = torch.einsum("cn,nij->cij", Bm, T)                             # [cell][md][m]
        va = torch.einsum("fce,cie->fci", Aft, Tc) / k_d                    # [f][cell][md]
        xa = xft + (Aft * Wc[None]).sum(2)                                  # [f][cell]
        v2f = ((Aft[:, :, :md] / k_d) ** 2).sum(2)
        v2a = (va ** 2).sum(2)
        out = []
        for f in range(nf):
            inf0 = infl[f]
            if math.isnan(ratio[f]):
                inf_f = inf0.expand(ncell)
            else:
                raw = torch.sqrt(v2f[f] / v2a[f]) * ratio[f] + 1.0 - ratio[f]
                inf_f = torch.where(raw < inf0, raw, inf0.expand(ncell))
            inf_f = torch.where(inf_f - 1.0 > 1e-6, inf_f, torch.ones_like(inf_f))
            out.append(inf_f[:, None] * va[f] + xa[f][:, None])
        return torch.stack(out)                                             # [f][cell][md]

    def J(Xa):
        return (((Xa - xt) ** 2) * boxt).sum()

    R0 = torch.tensor(P.obs_estd.astype(np.float64) ** 2)
    inn0 = torch.tensor(P.innov)
    iv = inn0.clone().requires_grad_(True)
    Rv = R0.clone().requires_grad_(True)
    fv = torch.tensor(infl0, dtype=torch.float64).requires_grad_(True)
    Xa = forward(iv, Rv, fv)
    Jr = J(Xa)
    Jr.backward()
    gy_ref, gR_ref, gi_ref = iv.grad.numpy(), Rv.grad.numpy(), fv.grad.numpy()
    # GPU: the calc and the ensemble, then the adjoint of dJ/dX^a
    eng = enkfg.Engine(m, nmem_dynamic=md, mode=mode, scheme=enkfg.SCHEME_DENKF, alpha=alpha, keep_fp64=True)
    eng.set_obstypes(P.obstypes)
    eng.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
    eng.set_S(P.HA, P.innov, standardised=False)
    eng.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    import synth_gpu
    dE = synth_gpu.DeviceArray.from_host(Ecell)
    eng.ens_attach(P.nlevels, P.stride, 0, 0, fields, dE.ptr)
    G = np.zeros((nf, ncell, m), np.float32)
    G[:, :, :md] = (2.0 * (Xa.detach() - xt) * boxt).numpy()
    dG = synth_gpu.DeviceArray.from_host(G)
    gy, gR, gi = eng.adjoint_ens(dG.ptr, nobs, nf)
    eng.close()
    ey = float(np.abs(gy - gy_ref).max() / np.abs(gy_ref).max())
    eR = float(np.abs(gR - gR_ref).max() / np.abs(gR_ref).max())
    ei = float(np.abs(gi - gi_ref).max() / max(np.abs(gi_ref).max(), 1e-300))
    print("  J %.6g  |dJ/dy| %.3g  |dJ/dR| %.3g  dJ/dinfl %s" % (Jr.item(), np.abs(gy_ref).max(), np.abs(gR_ref).max(),
                                                                np.array2string(gi_ref, precision=4)), flush=True)
    print("  error vs autograd: y %.2g  R %.2g  inflation %.2g" % (ey, eR, ei), flush=True)
    assert (gy_ref != 0).sum() > 50
    assert ey <= 1e-4 and eR <= 1e-4 and ei <= 1e-4
    # the T path matters: the mean-only gradients differ
    assert np.abs(gR_ref).max() > 0 and gi_ref[1] != 0.0


def test_adjoint_ens_c5_m96_against_autograd_and_finite_differences():
    """enkfg_adjoint_ens at the C5 size: 32 layers x 600 x 600, m = 96, 1M
    surface obs, LOCRAD 200 km, stride 3, DEnKF, capped inflation 1.05 (ratio
    1.0) on every field.  J = sum over a 100 x 100 box, the top 10 layers of
    temp and salt (20 fields) and the 96 members of (X^a - x^t)^2.  The
    reference is the float64 torch forward over the 1,225 stride nodes whose
    transforms reach the box (M, w, T = (I + M^-1)/2, their interpolation,
    v_a = A^f T^T, the capped inflation), differentiated by autograd, and
    central differences of it (h = 1e-6, relative for the variance) on the
    most sensitive observation.  Bound: 1e-4 of the largest gradient."""
    import time
    import synth_gpu
    print("case adjoint_ens C5: 32x600x600, m=96, 1M obs", flush=True)
    X = synth_gpu.make_tile(nj=600, ni=600, nk=32, nmem=96, n_sst=1_000_000, n_sla=0, n_prof=0, corr_km=150.0,
                            locrad=200.0, obs_std=(0.3, 0.04, 0.05), seed=5)
    m, s_ = X.nmem, X.stride
    nobs, ncell = X.nobs, X.ncell
    eng = enkfg.Engine(m, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF, keep_fp64=True)
    eng.set_obstypes(X.obstypes)
    eng.set_obs(X.obs_type, X.obs_status, X.obs_lon, X.obs_lat, X.obs_estd)
    eng.set_S(X.HA, X.innov, standardised=False)
    eng.calc_transforms(X.node_lon, X.node_lat, stride=s_)
    eng.ens_attach(X.nlevels, s_, 0, 0, X.fields, X.E.ptr)
    j0, i0, nb = 250, 250, 100
    fsel = list(range(10)) + list(range(X.nk, X.nk + 10))
    nf = len(fsel)
    rows = np.arange(j0, j0 + nb)
    cols = np.arange(i0, i0 + nb)
    cells = (rows[:, None] * X.ni + cols[None, :]).ravel()
    nc = len(cells)
    wet = np.stack([X.nlevels.ravel()[cells] > X.fields[f]["level"] for f in fsel]).astype(np.float64)
    Ef = np.stack([np.concatenate([X.E.read((f * ncell + j * X.ni + i0) * m, nb * m).reshape(nb, m) for j in rows])
                   for f in fsel])                                           # [f][cell][member] float32
    xf = Ef.astype(np.float64).mean(2)
    Af = (Ef.astype(np.float64) - xf[:, :, None]).astype(np.float32).astype(np.float64)
    # the four interpolation nodes and weights of each box cell
    jjs = sorted(set(rows // s_) | set(np.minimum(rows // s_ + 1, X.nj_s - 1)))
    iis = sorted(set(cols // s_) | set(np.minimum(cols // s_ + 1, X.ni_s - 1)))
    nodes = np.array([jj * X.ni_s + ii for jj in jjs for ii in iis])
    pos = {int(n): k for k, n in enumerate(nodes)}
    idx = np.zeros((nc, 4), np.int64)
    wgt = np.zeros((nc, 4))
    one = np.float32(1.0)
    for c, (j, i) in enumerate((j, i) for j in rows for i in cols):
        jj, sj, ii, si = j // s_, j % s_, i // s_, i % s_
        jj2, ii2 = min(jj + 1, X.nj_s - 1), min(ii + 1, X.ni_s - 1)
        wj2 = np.float32(sj) / np.float32(s_)
        wi2 = np.float32(si) / np.float32(s_)
        bj = (1.0, 0.0) if sj == 0 else (float(one - wj2), float(wj2))
        bi = (1.0, 0.0) if si == 0 else (float(one - wi2), float(wi2))
        q = 0
        for nj_, bjv in ((jj, bj[0]), (jj2, bj[1])):
            for ni_, biv in ((ii, bi[0]), (ii2, bi[1])):
                idx[c, q] = pos[nj_ * X.ni_s + ni_]
                wgt[c, q] = bjv * biv
                q += 1
    idxt = torch.tensor(idx)
    wgtt = torch.tensor(wgt)
    Aft = torch.tensor(Af)
    xft = torch.tensor(xf)
    wett = torch.tensor(wet)[:, :, None]
    jg, ig = np.meshgrid(rows, cols, indexing="ij")
    patt = (np.cos(2 * np.pi * ig / 50.0) * np.sin(2 * np.pi * jg / 70.0)).ravel()
    xt = torch.tensor(xf + np.array([0.3 if f < X.nk else 0.03 for f in fsel])[:, None] * patt[None, :])[:, :, None]
    ob = O.OracleObs(X.obstypes, X.obs_type, X.obs_lon, X.obs_lat, X.obs_estd)
    lists = [ob.findlocal(X.node_lon.ravel()[n], X.node_lat.ravel()[n]) for n in nodes]
    HA = torch.tensor(X.HA.astype(np.float64))
    sq = math.sqrt(m - 1)
    eye = torch.eye(m, dtype=torch.float64)

    def node_wT(k, innov, R):
        ids, lc = lists[k]
        if len(ids) == 0:
            return torch.zeros(m, dtype=torch.float64), eye
        it = torch.tensor(ids, dtype=torch.long)
        d = torch.sqrt(R[it]) * sq
        lct = torch.tensor(lc)
        S = HA[:, it].T / d[:, None] * lct[:, None]
        sv = innov[it] / d * lct
        Mi = torch.linalg.inv(eye + S.T @ S)
        return Mi @ (S.T @ sv), 0.5 * (eye + Mi)

    def analysis(W, T, infl):
        Wc = (W[idxt] * wgtt[:, :, None]).sum(1)                            # [cell][m]
        Tc = (T[idxt] * wgtt[:, :, None, None]).sum(1)                      # [cell][m][m]
        va = torch.einsum("fce,cie->fci", Aft, Tc)
        xa = xft + (Aft * Wc[None]).sum(2)
        v2f = (Aft ** 2).sum(2)
        v2a = (va ** 2).sum(2)
        raw = torch.sqrt(v2f / v2a)                                         # ratio 1.0
        inf = torch.where(raw < infl[:, None], raw, infl[:, None].expand_as(raw))
        inf = torch.where(inf - 1.0 > 1e-6, inf, torch.ones_like(inf))
        return inf[:, :, None] * va + xa[:, :, None]

    def J(Xa):
        return (((Xa - xt) ** 2) * wett).sum()

    R0 = torch.tensor(X.obs_estd.astype(np.float64) ** 2)
    inn0 = torch.tensor(X.innov)
    iv = inn0.clone().requires_grad_(True)
    Rv = R0.clone().requires_grad_(True)
    fv = torch.full((nf,), 1.05, dtype=torch.float64, requires_grad=True)
    t = time.perf_counter()
    WT = [node_wT(k, iv, Rv) for k in range(len(nodes))]
    W = torch.stack([a for a, _ in WT])
    T = torch.stack([b for _, b in WT])
    Xa = analysis(W, T, fv)
    Jr = J(Xa)
    Jr.backward()
    t_ref = time.perf_counter() - t
    gy_ref, gR_ref, gi_ref = iv.grad.numpy(), Rv.grad.numpy(), fv.grad.numpy()
    # GPU adjoint of dJ/dX^a
    G = (2.0 * (Xa.detach() - xt) * wett).numpy().astype(np.float32)      # [f][cell][m]
    dG = synth_gpu.DeviceArray((X.nfield, ncell, m), np.float32)
    dG.zero()
    for k, f in enumerate(fsel):
        for r, j in enumerate(rows):
            dG.write((f * ncell + j * X.ni + i0) * m, G[k, r * nb:(r + 1) * nb].ravel())
    t = time.perf_counter()
    gy, gR, gi = eng.adjoint_ens(dG.ptr, nobs, X.nfield)
    t_adj = time.perf_counter() - t
    eng.close()
    ey = float(np.abs(gy - gy_ref).max() / np.abs(gy_ref).max())
    eR = float(np.abs(gR - gR_ref).max() / np.abs(gR_ref).max())
    ei = float(np.abs(gi[fsel] - gi_ref).max() / np.abs(gi_ref).max())
    others = np.delete(gi, fsel)
    print("  %d nodes reach the box; J %.6g; dJ/dinfl of the box fields %.3g .. %.3g" % (
        len(nodes), Jr.item(), gi_ref.min(), gi_ref.max()), flush=True)
    print("  enkfg_adjoint_ens %.2f s (torch float64 reference %.1f s); error vs autograd: y %.2g  R %.2g  "
          "inflation %.2g" % (t_adj, t_ref, ey, eR, ei), flush=True)
    assert (gy_ref != 0).sum() > 10000 and np.abs(gi_ref).max() > 0
    assert ey <= 1e-4 and eR <= 1e-4 and ei <= 1e-4 and np.all(others == 0.0)
    # central differences on the most sensitive obs (only the nodes whose list
    # holds it change)
    with torch.no_grad():
        W0, T0 = W.detach().clone(), T.detach().clone()
        f0 = torch.full((nf,), 1.05, dtype=torch.float64)
        for name, base, gg in (("y", inn0, gy), ("R", R0, gR)):
            o = int(np.argmax(np.abs(gg)))
            ks = [k for k in range(len(nodes)) if o in set(lists[k][0].tolist())]
            h = 1e-6 if name == "y" else 1e-6 * float(base[o])
            Js = []
            for sgn in (1.0, -1.0):
                x = base.clone()
                x[o] += sgn * h
                Wp, Tp = W0.clone(), T0.clone()
                for k in ks:
                    Wp[k], Tp[k] = node_wT(k, x, R0) if name == "y" else node_wT(k, inn0, x)
                Js.append(J(analysis(Wp, Tp, f0)).item())
            fd = (Js[0] - Js[1]) / (2 * h)
            err = abs(fd - float(gg[o])) / float(np.abs(gg).max())
            print("  finite differences %s (obs %d, %d nodes): %.6g vs %.6g, error %.2g" % (
                name, o, len(ks), fd, float(gg[o]), err), flush=True)
            assert err <= 1e-4
