"""GPU parity of the batched SVD (many uneven matrices in one persistent
launch per class: columns longer than 128 rows (64 complex) in the wave
form, the rest in the lane-group form, on global memory or, when a matrix's
G and V fit the plan's LDS budget, resident in LDS).  Every matrix is checked against the oracle on
the same input with the same step table (sigma within 8 n eps, long-double
orthogonality and reconstruction) and against the plain GPU driver, which
performs the same operations per pair and must agree bit for bit."""
import numpy as np
import pytest

from oracle import vjo
from vecjac_amd import batched as BT
from vecjac_amd import svd as S
from vecjac_amd import workloads as W

pytestmark = pytest.mark.gpu
EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def _batch(nb, seed, nmax=48, dtype=np.float64):
    m, n = W.batch_shapes(nb, seed, nmin=2, nmax=nmax)
    m = np.maximum(m, n + (n & 1))
    m_, npad, offG, offV, offS, tot = BT.pack_shapes(m, n)
    G = BT.fill_gaussian(m, npad, n, offG, seed + 1, dtype)
    return m, n, npad, offG, offV, offS, tot, G


def _run(kind, m, npad, G, tot, waves=0, Gi=None, split=None, budget=None, wv=None, lanes=None, order=None,
         queue=None, V0=None, V0i=None, swp=999):
    import torch
    dt = torch.float64 if kind in "dz" else torch.float32
    plan = BT.BatchedSVD(m, npad, kind=kind, waves=waves, swp=swp)
    if V0 is not None:
        assert plan.vpreset(1)
    if order is not None:
        assert plan.order(order) == order
    if queue is not None:
        assert plan.queue(queue) == queue
    if budget is not None:
        plan.set(split if split is not None else (64 if kind in "zc" else 128), budget, wv, lanes)
    Gd = torch.from_numpy(G.copy()).cuda()
    Gid = torch.from_numpy(Gi.copy()).cuda() if Gi is not None else None
    V = torch.zeros(tot[1], dtype=dt, device="cuda") if V0 is None else torch.from_numpy(V0.copy()).cuda()
    Vi = None
    if Gi is not None:
        Vi = torch.zeros(tot[1], dtype=dt, device="cuda") if V0i is None else torch.from_numpy(V0i.copy()).cuda()
    eS = torch.zeros(tot[2], dtype=dt, device="cuda")
    fS = torch.zeros(tot[2], dtype=dt, device="cuda")
    info = torch.zeros(len(m), dtype=torch.int64, device="cuda")
    tr = torch.zeros(len(m), dtype=torch.int64, device="cuda")
    print(f"batched {kind} nb={len(m)} waves={plan.waves} order={plan.order()}", flush=True)
    rc = plan.run(Gd, V, eS, fS, info, tr, Gi=Gid, Vi=Vi)
    torch.cuda.synchronize()
    out = dict(rc=rc, G=Gd.cpu().numpy(), V=V.cpu().numpy(), eS=eS.cpu().numpy(), fS=fS.cpu().numpy(),
               info=info.cpu().numpy(), tr=tr.cpu().numpy(), evals=plan.evals().astype(np.int64),
               xfer=plan.traffic().astype(np.int64))
    plan.close()
    if Gi is not None:
        out["Gi"] = Gid.cpu().numpy()
        out["Vi"] = Vi.cpu().numpy()
    return out


@pytest.mark.parametrize("waves", [0, 1, 4, 8])
def test_batched_vs_oracle_f64(waves):
    nb = 48
    m, n, npad, offG, offV, offS, tot, G = _batch(nb, 100 + waves)
    r = _run("d", m, npad, G, tot, waves=waves)
    assert r["rc"] == 0
    for b in range(nb):
        mm, nn = int(m[b]), int(npad[b])
        A = G[offG[b]: offG[b] + mm * nn].reshape(nn, mm)
        js, stp = S.pjs(6, nn)
        sg = S.sigma(r["eS"][offS[b]: offS[b] + nn], r["fS"][offS[b]: offS[b] + nn])
        for mode in (1, 0):  # the GPU's norm and dot; the reference's own
            Go = vjo.aligned((nn, mm), np.float64)
            Go[...] = A
            with vjo.norm_mode(mode):
                rc, eS, fS, V, _, st = vjo.vjsvd_core("d", Go, js=js, stp=stp, swp=999)
            assert rc >= 0 and abs(int(r["info"][b]) - rc) <= 1, (b, mode, r["info"][b], rc)
            so = S.sigma(eS, fS)
            sg_s, so_s = np.sort(sg), np.sort(so)
            nz = so_s > 0
            assert np.max(np.abs(sg_s[nz] - so_s[nz]) / so_s[nz]) <= 8 * nn * EPS64, (b, mode)
            assert np.all(sg_s[~nz] == 0)
        U = r["G"][offG[b]: offG[b] + mm * nn].reshape(nn, mm)[sg > 0].astype(np.longdouble)
        Vb = r["V"][offV[b]: offV[b] + nn * nn].reshape(nn, nn).astype(np.longdouble)
        assert np.max(np.abs(U @ U.T - np.eye(U.shape[0], dtype=np.longdouble))) <= 8 * nn * EPS64
        assert np.max(np.abs(Vb @ Vb.T - np.eye(nn, dtype=np.longdouble))) <= 8 * nn * EPS64


def test_batched_equals_plain_bitwise():
    import torch
    nb = 24
    m, n, npad, offG, offV, offS, tot, G = _batch(nb, 7)
    r = _run("d", m, npad, G, tot, waves=2)
    same = 0
    for b in range(nb):
        mm, nn = int(m[b]), int(npad[b])
        A = G[offG[b]: offG[b] + mm * nn].reshape(nn, mm)
        js, stp = S.pjs(6, nn)
        Gd = torch.from_numpy(A.copy()).cuda()
        Vd = torch.empty((nn, nn), dtype=torch.float64, device="cuda")
        eS = torch.empty(nn, dtype=torch.float64, device="cuda")
        fS = torch.empty(nn, dtype=torch.float64, device="cuda")
        jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
        rc, _ = S.device_vjsvd("d", Gd, Vd, eS, fS, jsd, stp)
        assert rc == int(r["info"][b])
        # U columns of zero singular values (padded widths) are NaN in both
        ok = (np.array_equal(Gd.cpu().numpy().reshape(-1), r["G"][offG[b]: offG[b] + mm * nn], equal_nan=True) and
              np.array_equal(eS.cpu().numpy(), r["eS"][offS[b]: offS[b] + nn]) and
              np.array_equal(fS.cpu().numpy(), r["fS"][offS[b]: offS[b] + nn]))
        if not ok:
            print(f"matrix {b} (m={mm}, n={nn}) differs")
        same += ok
    assert same == nb


@pytest.mark.parametrize("kind", ["s", "z", "c"])
def test_batched_other_precisions(kind):
    nb = 24
    dt = np.float64 if kind in "dz" else np.float32
    eps = EPS64 if dt == np.float64 else EPS32
    m, n, npad, offG, offV, offS, tot, G = _batch(nb, 55, nmax=32, dtype=dt)
    Gi = BT.fill_gaussian(m, npad, n, offG, 56, dt) if kind in "zc" else None
    r = _run(kind, m, npad, G, tot, Gi=Gi)
    assert r["rc"] == 0 and np.all(r["info"] >= 0)
    for b in range(nb):
        mm, nn = int(m[b]), int(npad[b])
        A = G[offG[b]: offG[b] + mm * nn].reshape(nn, mm).astype(np.complex128)
        if Gi is not None:
            A = A + 1j * Gi[offG[b]: offG[b] + mm * nn].reshape(nn, mm)
        sv = np.linalg.svd(A.T, compute_uv=False)
        sg = np.sort(S.sigma(r["eS"][offS[b]: offS[b] + nn], r["fS"][offS[b]: offS[b] + nn]))[::-1]
        nz = sv > 1e-5 * sv[0]
        assert np.max(np.abs(sg[nz] - sv[nz]) / sv[nz]) <= 16 * nn * eps


def test_batched_nonfinite_and_zero_members():
    m = np.array([16, 24, 40, 8], dtype=np.int64)
    n = np.array([8, 8, 16, 8], dtype=np.int64)
    _, npad, offG, offV, offS, tot = BT.pack_shapes(m, n)
    G = BT.fill_gaussian(m, npad, n, offG, 3)
    G[offG[1] + 5] = np.nan
    G[offG[3]: offG[3] + 64] = 0.0
    r = _run("d", m, npad, G, tot)
    assert r["info"][1] == -80 and r["info"][3] == 0
    assert r["info"][0] >= 1 and r["info"][2] >= 1


@pytest.mark.parametrize("order", [6, 4])
def test_batched_wide_rows_vs_plain_and_oracle(order):
    """Widths 64..128 and up to 512 rows: the register footprints of 4 and 8
    row chunks per lane (block rounds of four columns with and without
    prefetch under ordering 6; two pairs per wave under ordering 4) against
    the plain driver run with the same table bit for bit, and two members
    against the oracle."""
    import torch
    g = W.rng(321)
    n = np.array([128, 96, 64, 128, 80, 112], dtype=np.int64)
    m = np.array([512, 300, 256, 130, 500, 448], dtype=np.int64)
    _, npad, offG, offV, offS, tot = BT.pack_shapes(m, n)
    G = BT.fill_gaussian(m, npad, n, offG, 322)
    r = _run("d", m, npad, G, tot, order=order)
    assert r["rc"] == 0 and np.all(r["info"] > 0)
    for b in range(len(m)):
        mm, nn = int(m[b]), int(npad[b])
        A = G[offG[b]: offG[b] + mm * nn].reshape(nn, mm)
        js, stp = S.pjs(order, nn)
        Gd = torch.from_numpy(A.copy()).cuda()
        Vd = torch.empty((nn, nn), dtype=torch.float64, device="cuda")
        eS = torch.empty(nn, dtype=torch.float64, device="cuda")
        fS = torch.empty(nn, dtype=torch.float64, device="cuda")
        jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
        print(f"plain m={mm} n={nn}", flush=True)
        rc, _ = S.device_vjsvd("d", Gd, Vd, eS, fS, jsd, stp)
        assert rc == int(r["info"][b])
        assert np.array_equal(Gd.cpu().numpy().reshape(-1), r["G"][offG[b]: offG[b] + mm * nn], equal_nan=True)
        assert np.array_equal(Vd.cpu().numpy().reshape(-1), r["V"][offV[b]: offV[b] + nn * nn])
        assert np.array_equal(eS.cpu().numpy(), r["eS"][offS[b]: offS[b] + nn])
        assert np.array_equal(fS.cpu().numpy(), r["fS"][offS[b]: offS[b] + nn])
    for b in (1, 3):
        mm, nn = int(m[b]), int(npad[b])
        js, stp = S.pjs(order, nn)
        Go = vjo.aligned((nn, mm), np.float64)
        Go[...] = G[offG[b]: offG[b] + mm * nn].reshape(nn, mm)
        with vjo.norm_mode(1):
            rc, eSo, fSo, Vo, _, st = vjo.vjsvd_core("d", Go, js=js, stp=stp, swp=999)
        assert rc == int(r["info"][b])
        assert np.array_equal(eSo, r["eS"][offS[b]: offS[b] + nn])
        assert np.array_equal(fSo, r["fS"][offS[b]: offS[b] + nn])


def test_batched_graph_replay():
    """The run captured in a HIP graph (its queue is reset by a fill kernel,
    not a memset node) replays to the eager result, replay after replay, with
    eager work on the same buffers in between."""
    import torch
    m, n, npad, offG, offV, offS, tot, G = _batch(40, 808, nmax=64)
    plan = BT.BatchedSVD(m, npad, kind="d")
    G0 = torch.from_numpy(G).cuda()
    bufs = lambda: (torch.empty_like(G0), torch.zeros(tot[1], dtype=torch.float64, device="cuda"),
                    torch.zeros(tot[2], dtype=torch.float64, device="cuda"),
                    torch.zeros(tot[2], dtype=torch.float64, device="cuda"),
                    torch.zeros(len(m), dtype=torch.int64, device="cuda"))
    Ge, Ve, eSe, fSe, ie = bufs()
    Ge.copy_(G0)
    plan.run(Ge, Ve, eSe, fSe, ie)
    torch.cuda.synchronize()
    ref = [t.cpu().numpy() for t in (Ge, Ve, eSe, fSe, ie)]
    Gg, Vg, eSg, fSg, ig = bufs()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        Gg.copy_(G0)
        plan.run(Gg, Vg, eSg, fSg, ig)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        Gg.copy_(G0)
        plan.run(Gg, Vg, eSg, fSg, ig)
    for rep in range(3):
        for t in (Gg, Vg, eSg, fSg):
            t.fill_(float(rep) + 0.5)
        ig.fill_(-7)
        torch.cuda.synchronize()
        print(f"graph replay {rep}", flush=True)
        graph.replay()
        torch.cuda.synchronize()
        got = [t.cpu().numpy() for t in (Gg, Vg, eSg, fSg, ig)]
        for a, b in zip(got, ref):
            assert np.array_equal(a, b, equal_nan=True), rep
    plan.close()


def test_batched_autotune_keeps_results():
    """The autotuner leaves G untouched, picks a launch shape, and the tuned
    plan computes the same bits as any other shape."""
    import torch
    m, n, npad, offG, offV, offS, tot, G = _batch(32, 909, nmax=64)
    base = _run("d", m, npad, G, tot, waves=4)
    plan = BT.BatchedSVD(m, npad, kind="d")
    G0 = torch.from_numpy(G).cuda()
    V = torch.zeros(tot[1], dtype=torch.float64, device="cuda")
    eS = torch.zeros(tot[2], dtype=torch.float64, device="cuda")
    fS = torch.zeros(tot[2], dtype=torch.float64, device="cuda")
    info = torch.zeros(len(m), dtype=torch.int64, device="cuda")
    times = plan.autotune(G0, V, eS, fS, info)
    print(f"autotune {times} -> {plan.waves} waves", flush=True)
    assert plan.waves in times and len(times) >= 3
    assert np.array_equal(G0.cpu().numpy(), G)
    Gd = G0.clone()
    plan.run(Gd, V, eS, fS, info)
    torch.cuda.synchronize()
    assert np.array_equal(Gd.cpu().numpy(), base["G"], equal_nan=True)
    assert np.array_equal(eS.cpu().numpy(), base["eS"]) and np.array_equal(fS.cpu().numpy(), base["fS"])
    plan.close()


@pytest.mark.parametrize("kind,order", [("d", 6), ("d", 4), ("z", 6), ("z", 4)])
def test_short_form_vs_plain_bitwise(kind, order):
    """Members on both sides of the lane-group bounds (4 chunks of 16 lanes up
    to 64 rows, 8 chunks up to 128; complex up to 64 rows) and just past them
    in the wave form (block rounds for real data under ordering 6), against
    the plain driver run with the same table bit for bit."""
    import torch
    m = np.array([8, 16, 40, 63, 64, 65, 96, 127, 128, 129, 200], dtype=np.int64)
    n = np.array([8, 10, 40, 30, 64, 34, 48, 64, 128, 66, 100], dtype=np.int64)
    _, npad, offG, offV, offS, tot = BT.pack_shapes(m, n)
    G = BT.fill_gaussian(m, npad, n, offG, 431)
    Gi = BT.fill_gaussian(m, npad, n, offG, 432) if kind == "z" else None
    r = _run(kind, m, npad, G, tot, Gi=Gi, order=order)
    assert r["rc"] == 0 and np.all(r["info"] > 0)
    for b in range(len(m)):
        mm, nn = int(m[b]), int(npad[b])
        js, stp = S.pjs(order, nn)
        Gd = torch.from_numpy(G[offG[b]: offG[b] + mm * nn].reshape(nn, mm).copy()).cuda()
        Vd = torch.empty((nn, nn), dtype=torch.float64, device="cuda")
        Gid = Vid = None
        if kind == "z":
            Gid = torch.from_numpy(Gi[offG[b]: offG[b] + mm * nn].reshape(nn, mm).copy()).cuda()
            Vid = torch.empty((nn, nn), dtype=torch.float64, device="cuda")
        eS = torch.empty(nn, dtype=torch.float64, device="cuda")
        fS = torch.empty(nn, dtype=torch.float64, device="cuda")
        jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
        print(f"plain {kind} m={mm} n={nn}", flush=True)
        rc, _ = S.device_vjsvd(kind, Gd, Vd, eS, fS, jsd, stp, Gi=Gid, Vi=Vid)
        assert rc == int(r["info"][b]), (b, rc, r["info"][b])
        sl = slice(offG[b], offG[b] + mm * nn)
        assert np.array_equal(Gd.cpu().numpy().reshape(-1), r["G"][sl], equal_nan=True), b
        assert np.array_equal(Vd.cpu().numpy().reshape(-1), r["V"][offV[b]: offV[b] + nn * nn]), b
        if kind == "z":
            assert np.array_equal(Gid.cpu().numpy().reshape(-1), r["Gi"][sl], equal_nan=True), b
            assert np.array_equal(Vid.cpu().numpy().reshape(-1), r["Vi"][offV[b]: offV[b] + nn * nn]), b
        assert np.array_equal(eS.cpu().numpy(), r["eS"][offS[b]: offS[b] + nn]), b
        assert np.array_equal(fS.cpu().numpy(), r["fS"][offS[b]: offS[b] + nn]), b


def test_pair_evaluation_counts():
    """vjh_bplan_evals: every matrix evaluates at most sweeps x n steps x n/2
    pairs (fewer when unchanged pairs are not evaluated again), at least one
    full sweep, and never fewer pairs than it transformed."""
    import torch
    m, n, npad, offG, offV, offS, tot, G = _batch(32, 919, nmax=64)
    plan = BT.BatchedSVD(m, npad, kind="d")
    Gd = torch.from_numpy(G.copy()).cuda()
    V = torch.zeros(tot[1], dtype=torch.float64, device="cuda")
    eS = torch.zeros(tot[2], dtype=torch.float64, device="cuda")
    fS = torch.zeros(tot[2], dtype=torch.float64, device="cuda")
    info = torch.zeros(len(m), dtype=torch.int64, device="cuda")
    tr = torch.zeros(len(m), dtype=torch.int64, device="cuda")
    print("batched evals", flush=True)
    plan.run(Gd, V, eS, fS, info, tr)
    torch.cuda.synchronize()
    ev = plan.evals().astype(np.int64)
    plan.close()
    info, tr = info.cpu().numpy(), tr.cpu().numpy()
    nn = npad.astype(np.int64)
    slots = (info + 1) * nn * (nn // 2)
    assert np.all(info > 0)
    assert np.all(ev <= slots) and np.all(ev >= nn * (nn // 2)) and np.all(ev >= tr)
    print(f"evaluated {ev.sum()} of {slots.sum()} pair slots")


@pytest.mark.parametrize("kind", ["d", "z", "s", "c"])
def test_lds_resident_form_bitwise(kind):
    """The lane-group form resident in LDS computes the same bits as on
    global memory, and so do its narrower groups (8 and 4 lanes per pair):
    one batch (long members in the wave form, short ones of every width) run
    with no LDS budget, with the largest one and with the default, at 16, 8
    and 4 lanes per pair and several waves per workgroup; every output, the
    sweep counts and the transformation counts agree.  Also at a split of
    half the short-row bound, which moves members between the launches."""
    dt = np.float64 if kind in "dz" else np.float32
    m = np.array([8, 16, 40, 63, 64, 96, 127, 128, 200, 24, 48, 32, 64, 100], dtype=np.int64)
    n = np.array([8, 10, 40, 30, 64, 34, 48, 64, 100, 24, 16, 32, 32, 90], dtype=np.int64)
    if kind in "zc":
        m = np.minimum(m, 64)
        m = np.maximum(m, n + (n & 1))  # members wider than 64 stay in the wave form
    _, npad, offG, offV, offS, tot = BT.pack_shapes(m, n)
    G = BT.fill_gaussian(m, npad, n, offG, 771, dt)
    Gi = BT.fill_gaussian(m, npad, n, offG, 772, dt) if kind in "zc" else None
    sr = 64 if kind in "zc" else 128
    runs = {}
    shapes = [(sr, 0, None, None), (sr, 152 * 1024, None, None), (sr, None, None, None),
              (sr // 2, 152 * 1024, None, None),
              # 8 lanes per pair (8 pairs per wave), on global memory and in LDS
              (sr, 0, (4, 2, 2), (16, 8, 8)), (sr, 152 * 1024, (4, 4, 1), (16, 8, 8))]
    if kind in "ds":  # 4 lanes per pair (16 pairs per wave), real rows up to 64
        shapes += [(64, 0, (4, 1, 2), (16, 4, 4)), (64, 152 * 1024, (4, 2, 1), (16, 4, 4))]
    for split, budget, wv, lanes in shapes:
        r = _run(kind, m, npad, G, tot, Gi=Gi, split=split, budget=budget, wv=wv, lanes=lanes)
        assert r["rc"] == 0 and np.all(r["info"] > 0), (split, budget, r["info"])
        runs[(split, budget, wv, lanes)] = r
    ref = runs[(sr, 0, None, None)]
    for key, r in runs.items():
        for f in ("G", "V", "eS", "fS", "info", "tr") + (("Gi", "Vi") if kind in "zc" else ()):
            assert np.array_equal(r[f], ref[f], equal_nan=True), (key, f)


def test_block_rounds_halve_column_traffic():
    """vjh_bplan_traffic: the long-column members move about half the G and
    V elements per pair evaluation in block rounds (ordering 6: four columns
    loaded and stored once for four evaluations) that they move in single
    steps (ordering 4: two columns per evaluation); both counts cover at
    least the set-up scan, V := I and one sweep of loads."""
    m = np.array([160, 256, 300, 512, 400, 200], dtype=np.int64)
    n = np.array([32, 64, 96, 128, 80, 40], dtype=np.int64)
    _, npad, offG, offV, offS, tot = BT.pack_shapes(m, n)
    G = BT.fill_gaussian(m, npad, n, offG, 515)
    per = {}
    for order in (6, 4):
        r = _run("d", m, npad, G, tot, order=order)
        assert r["rc"] == 0 and np.all(r["info"] > 0)
        nn = npad.astype(np.int64)
        floor = m * nn + nn * nn + (nn // 2) * nn * (m + nn)
        assert np.all(r["xfer"] >= floor), (order, r["xfer"], floor)
        per[order] = r["xfer"] / r["evals"]
    ratio = per[6] / per[4]
    print("traffic per evaluation, ordering 6 / ordering 4:", np.round(ratio, 3))
    assert np.all(ratio < 0.75), ratio


def test_queue_models_same_bits():
    """The processing order's cost model (vjh_bplan_queue) changes only the
    order in which workgroups take matrices: every output is the same."""
    m, n, npad, offG, offV, offS, tot, G = _batch(40, 313, nmax=96)
    runs = [_run("d", m, npad, G, tot, queue=q) for q in (0, 1, 2)]
    for r in runs[1:]:
        for f in ("G", "V", "eS", "fS", "info", "tr"):
            assert np.array_equal(r[f], runs[0][f], equal_nan=True), f


@pytest.mark.parametrize("kind", ["d", "z"])
def test_preset_v_vs_plain_bitwise(kind):
    """A preset V (vjh_bplan_vpreset, the batched iwork[0] = 1): each member
    against the plain driver run with vpreset on the same G, V and table,
    bit for bit: long and short members, both sweep-limited first passes'
    outputs resumed and random orthogonal starting matrices scaled by 2^300
    (which takes the initial and tracked V rescalings); a zero V fails with
    -13 (-17), an infinite entry with -92 (-104)."""
    import torch
    m = np.array([200, 96, 64, 300, 48], dtype=np.int64)
    n = np.array([64, 40, 32, 96, 48], dtype=np.int64)
    _, npad, offG, offV, offS, tot = BT.pack_shapes(m, n)
    G = BT.fill_gaussian(m, npad, n, offG, 611)
    Gi = BT.fill_gaussian(m, npad, n, offG, 612) if kind == "z" else None
    g = W.rng(613)
    V0 = np.zeros(tot[1])
    V0i = np.zeros(tot[1]) if kind == "z" else None
    for b in range(len(m)):
        nn = int(npad[b])
        Q = W.haar_unitary(nn, nn, g, kind == "z") * 2.0 ** 300
        V0[offV[b]: offV[b] + nn * nn] = Q.real.T.reshape(-1)
        if kind == "z":
            V0i[offV[b]: offV[b] + nn * nn] = Q.imag.T.reshape(-1)
    V0[offV[2]: offV[2] + int(npad[2]) ** 2] = 0.0          # member 2: zero V
    if kind == "z":
        V0i[offV[2]: offV[2] + int(npad[2]) ** 2] = 0.0
    V0[offV[4] + 7] = np.inf                                # member 4: infinite entry
    r = _run(kind, m, npad, G, tot, Gi=Gi, V0=V0, V0i=V0i)
    assert r["info"][2] == (-17 if kind == "z" else -13)
    assert r["info"][4] == (-104 if kind == "z" else -92)
    for b in (0, 1, 3):
        mm, nn = int(m[b]), int(npad[b])
        js, stp = S.pjs(6, nn)
        sl, sv = slice(offG[b], offG[b] + mm * nn), slice(offV[b], offV[b] + nn * nn)
        Gd = torch.from_numpy(G[sl].reshape(nn, mm).copy()).cuda()
        Vd = torch.from_numpy(V0[sv].reshape(nn, nn).copy()).cuda()
        Gid = torch.from_numpy(Gi[sl].reshape(nn, mm).copy()).cuda() if kind == "z" else None
        Vid = torch.from_numpy(V0i[sv].reshape(nn, nn).copy()).cuda() if kind == "z" else None
        eS = torch.empty(nn, dtype=torch.float64, device="cuda")
        fS = torch.empty(nn, dtype=torch.float64, device="cuda")
        jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
        print(f"plain {kind} vpreset m={mm} n={nn}", flush=True)
        rc, _ = S.device_vjsvd(kind, Gd, Vd, eS, fS, jsd, stp, Gi=Gid, Vi=Vid, vpreset=True)
        assert rc == int(r["info"][b]) and rc > 0, (b, rc, r["info"][b])
        assert np.array_equal(Gd.cpu().numpy().reshape(-1), r["G"][sl], equal_nan=True), b
        assert np.array_equal(Vd.cpu().numpy().reshape(-1), r["V"][sv]), b
        if kind == "z":
            assert np.array_equal(Gid.cpu().numpy().reshape(-1), r["Gi"][sl], equal_nan=True), b
            assert np.array_equal(Vid.cpu().numpy().reshape(-1), r["Vi"][sv]), b
        assert np.array_equal(eS.cpu().numpy(), r["eS"][offS[b]: offS[b] + nn]), b
        assert np.array_equal(fS.cpu().numpy(), r["fS"][offS[b]: offS[b] + nn]), b


def test_preset_v_resume_equals_plain_resume():
    """A sweep-limited batched run (2 sweeps) resumed with its own working G
    and V as the preset equals the plain driver resumed the same way, bit
    for bit, for every member."""
    import torch
    m, n, npad, offG, offV, offS, tot, G = _batch(12, 717, nmax=80)
    r1 = _run("d", m, npad, G, tot, swp=2)
    assert np.all(r1["info"] == 2)
    r2 = _run("d", m, npad, r1["G"], tot, V0=r1["V"])
    for b in range(len(m)):
        mm, nn = int(m[b]), int(npad[b])
        js, stp = S.pjs(6, nn)
        sl, sv = slice(offG[b], offG[b] + mm * nn), slice(offV[b], offV[b] + nn * nn)
        Gd = torch.from_numpy(G[sl].reshape(nn, mm).copy()).cuda()
        Vd = torch.empty((nn, nn), dtype=torch.float64, device="cuda")
        eS = torch.empty(nn, dtype=torch.float64, device="cuda")
        fS = torch.empty(nn, dtype=torch.float64, device="cuda")
        jsd = torch.from_numpy(js.reshape(-1).view(np.int32)).cuda()
        rc1, _ = S.device_vjsvd("d", Gd, Vd, eS, fS, jsd, stp, swp=2)
        assert rc1 == 2
        rc2, _ = S.device_vjsvd("d", Gd, Vd, eS, fS, jsd, stp, vpreset=True)
        assert rc2 == int(r2["info"][b]) and rc2 >= 0, (b, rc2, r2["info"][b])
        assert np.array_equal(Gd.cpu().numpy().reshape(-1), r2["G"][sl], equal_nan=True), b
        assert np.array_equal(Vd.cpu().numpy().reshape(-1), r2["V"][sv]), b
        assert np.array_equal(eS.cpu().numpy(), r2["eS"][offS[b]: offS[b] + nn]), b
