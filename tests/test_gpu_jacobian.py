"""GPU checks of the differentiable form (umcx_amd.jacobian, umcx_session_replay).

* per-voxel properties equal to the labels' media change nothing (bitwise);
* the replay kernel re-traces each detected photon to the same final weight (bitwise);
* per medium, the replayed d log I / d mua equals the perturbation-MC gradient computed on the CPU
  from the partial path lengths the *reference binary* saved for the same photons;
* central finite differences under common random numbers (relative step 1e-3) on 20 random
  voxels agree with the replayed d log I / d mua;
* the mus derivative summed over the background agrees with central differences of the
  background's mus within the Monte Carlo noise measured from independent seeds, and the same
  estimator without its n/mus term does not;
* the torch.autograd.Function's gradient equals the model's vector-Jacobian product.
"""
import copy
import json
import sys

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U
from umcx_amd import jacobian as J

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def srt(r):
    return r[np.lexsort(r.T[::-1])] if r.size else r


def label_props(cfg):
    lab = U.domain(cfg) & 0x7FFFFFFF
    unit = np.float32(cfg["Domain"].get("LengthUnit", 1))
    med = np.array([[m["mua"], m["mus"]] if isinstance(m, dict) else m[:2] for m in cfg["Domain"]["Media"]],
                   dtype=np.float32)
    return med[lab, 0].copy(), med[lab, 1].copy(), lab


def test_voxel_props_equal_to_labels_are_bitwise_neutral():
    g = O.golden("c2")
    say("voxel props neutral")
    mua, mus, _ = label_props(g["cfg"])
    with U.Simulation(g["cfg"], tally_mode=0) as a:
        a.run()
        ra, ea, fa = srt(a.detp()[0]), a.energies(), a.fluence()
    with U.Simulation(g["cfg"], tally_mode=0) as b:
        from umcx_amd._native import lib
        assert lib().umcx_session_set_voxel_props(b._h, mua.ctypes.data, mus.ctypes.data, 0) == 0
        b.run()
        rb, eb, fb = srt(b.detp()[0]), b.energies(), b.fluence()
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert ea[1] == eb[1] and ea[0] == pytest.approx(eb[0], rel=1e-12)
    # float32 atomics in another order: the fluence tolerance of the parity bar (1e-4 rel L2)
    assert np.linalg.norm(fa - fb) <= 1e-4 * np.linalg.norm(fa)


def c4_model(nsrc_side=2, nphoton=200000, cfg=None):
    g = O.golden("c4src")
    cfg = copy.deepcopy(cfg or g["cfg"])
    srcs = J.grid_sources(64, 64, nsrc_side, 16.5) if nsrc_side else [((24.5, 24.5, 0.0), (0.0, 0.0, 1.0))]
    return cfg, J.MultiSourceModel(cfg, srcs, nphoton)


def test_replay_reproduces_detected_weights():
    cfg, m = c4_model(nsrc_side=0, nphoton=20000)
    say("replay weights")
    m.forward()
    ids, det, w = m.records[0]
    assert len(ids) == int(O.golden("c4src")["detected"])
    rw = m.replay_weights(0)
    assert np.array_equal(rw.view(np.uint32), w.view(np.uint32))
    m.close()


def test_jacobian_matches_reference_partial_paths():
    """Same photons as the reference run: sum over each medium's voxels of d log I_d / d mua_v
    must equal -sum_p w_p L_pm / I_d built from the reference's saved partial paths."""
    g = O.golden("c4src")
    cfg, m = c4_model(nsrc_side=0, nphoton=20000)
    say("jacobian vs reference ppath")
    m.forward()
    Jt = m.jacobian().cpu().numpy()[:, :, 0]           # (D, nvox) d log I / d mua
    _, _, lab = label_props(cfg)
    lab = lab.reshape(-1)
    unit = float(cfg["Domain"]["LengthUnit"])
    media = cfg["Domain"]["Media"]
    rows = g["detp"].astype(np.float64)                 # [detid, ppath m0, m1, m2] from the reference
    det = rows[:, 0].astype(int) - 1
    pp = rows[:, 1:4] * unit                            # mm
    mua = np.array([media[1]["mua"], media[2]["mua"], 0.0], dtype=np.float64)   # column k is medium k+1
    w = np.exp(-(pp * mua).sum(axis=1))
    checked = 0
    for d in range(m.ndet):
        sel = det == d
        if sel.sum() < 5:
            continue
        I = w[sel].sum()
        for med in (1, 2):
            cpu = -(w[sel] * pp[sel, med - 1]).sum() / I
            gpu = Jt[d, lab == med].astype(np.float64).sum()
            assert gpu == pytest.approx(cpu, rel=2e-4, abs=1e-9), (d, med, gpu, cpu)
            checked += 1
    assert checked >= 8
    m.close()


def test_mua_central_differences_under_common_random_numbers():
    cfg, m = c4_model(nsrc_side=2, nphoton=400000)
    say("FD mua CRN")
    mua0, mus0, lab = label_props(cfg)
    m.forward(mua0, mus0)
    Jm = m.jacobian().cpu().numpy()[:, :, 0]          # (S*D, nvox)
    base = m.I.copy()
    sens = np.abs(Jm).max(axis=0)
    rng = np.random.default_rng(3)
    cand = np.flatnonzero(sens > 0.05 * sens.max())
    vox = rng.choice(cand, size=20, replace=False)
    h = 1e-3
    errs = []
    for v in vox:
        out = []
        for sgn in (1, -1):
            mua = mua0.copy()
            mua.reshape(-1)[v] *= (1 + sgn * h)
            out.append(m.forward(mua, mus0).reshape(-1))
        fd = (out[0] - out[1]) / (2 * h * float(mua0.reshape(-1)[v]))
        an = Jm[:, v]
        ok = np.abs(an) > 0.1 * np.abs(an).max()
        errs.append(np.linalg.norm(fd[ok] - an[ok]) / np.linalg.norm(an[ok]))
    say(f"FD relative errors: {np.round(errs, 5)}")
    # a 1e-3 change of one voxel's mua moves each crossing photon's float32 weight by ~1e-5
    # relative, against ~1e-6 of rounding noise per photon: the FD estimate carries that noise
    assert np.max(errs) < 3e-2
    assert np.median(errs) < 1e-2
    m.close()


def test_mus_derivative_within_monte_carlo_noise():
    """d log I / d mus summed over the background medium's voxels vs central differences of the
    background mus (relative step 1e-2, common random numbers), over 4 independent seeds."""
    g = O.golden("c4src")
    fds, pms, wrong = [], [], []
    h = 1e-2
    for seed in (11, 12, 13, 14):
        cfg = copy.deepcopy(g["cfg"])
        cfg["Session"]["RNGSeed"] = seed
        cfg, m = c4_model(nsrc_side=0, nphoton=1_000_000, cfg=cfg)
        mua0, mus0, lab = label_props(cfg)
        bg = (lab == 1).reshape(-1)
        m.forward(mua0, mus0)
        Jall = m.jacobian().cpu().numpy()
        pms.append(Jall[:, bg, 1].astype(np.float64).sum(axis=1))
        # the same estimator without its n/mus term is -sum w L / I = the summed mua derivative
        wrong.append(Jall[:, bg, 0].astype(np.float64).sum(axis=1))
        out = []
        for sgn in (1, -1):
            mus = mus0.copy().reshape(-1)
            mus[bg] *= (1 + sgn * h)
            out.append(m.forward(mua0, mus.reshape(mus0.shape)).reshape(-1))
        fds.append((out[0] - out[1]) / (2 * h * float(mus0.reshape(-1)[bg][0])))
        m.close()
    fds, pms = np.array(fds), np.array(pms)
    say(f"mus: pMC {np.round(pms.mean(0)[:6], 4)}  FD {np.round(fds.mean(0)[:6], 4)}")
    se = np.sqrt(fds.var(0, ddof=1) / 4 + pms.var(0, ddof=1) / 4)
    ok = np.isfinite(fds.mean(0)) & (se > 0)
    assert ok.sum() >= 8
    z = np.abs(fds.mean(0) - pms.mean(0))[ok] / se[ok]
    say(f"mus z-scores: {np.round(z, 2)}")
    assert np.median(z) < 2.0 and z.max() < 5.0
    # dropping the n/mus term (a plausible mistake) lands far outside that band
    zw = np.abs(fds.mean(0) - np.array(wrong).mean(0))[ok] / se[ok]
    assert np.median(zw) > 10


def test_autograd_function_matches_vjp():
    torch = pytest.importorskip("torch")
    cfg, m = c4_model(nsrc_side=2, nphoton=200000)
    say("autograd")
    mua0, mus0, _ = label_props(cfg)
    mua = torch.tensor(mua0, device="cuda", requires_grad=True)
    mus = torch.tensor(mus0, device="cuda", requires_grad=True)
    logI = J.log_intensity(mua, mus, m)
    wts = torch.randn_like(logI)
    (logI * wts).sum().backward()
    Jf = m.jacobian()                                   # (S*D, nvox, 2)
    ref_mua = (wts.reshape(-1, 1) * Jf[:, :, 0]).sum(0).reshape(mua.shape)
    ref_mus = (wts.reshape(-1, 1) * Jf[:, :, 1]).sum(0).reshape(mus.shape)
    # both sides are float32 atomic sums in different orders; the mus terms (n/mus - L) have both
    # signs, so an entry's rounding error scales with the weighted sum of the slices' magnitudes,
    # not with the entry itself
    for k, (got, ref) in enumerate(((mua.grad, ref_mua), (mus.grad, ref_mus))):
        scale = (wts.abs().reshape(-1, 1) * Jf[:, :, k].abs()).sum(0).reshape(ref.shape)
        bound = 2e-4 * scale + 1e-6 * float(ref.abs().max())
        excess = float(((got - ref).abs() - bound).max())
        assert excess <= 0.0, (k, excess)
    m.close()


def test_one_launch_replay_equals_per_source_replays():
    """umcx_session_replay_sources (every source's records in one launch) adds the same
    gradients as one umcx_session_replay per source (float summation order aside)."""
    import torch
    cfg, m = c4_model(nsrc_side=2, nphoton=100000)
    say("one-launch replay vs per-source replays")
    m.forward()
    Jall = m.jacobian()
    S, D = len(m.sources), m.ndet
    J2 = torch.zeros((2, S * D, m.nvox), dtype=torch.float32, device="cuda")
    for s in range(S):
        ids, det, w = m.records[s]
        Isd = m.I[s, det]
        coef = np.where(Isd > 0, w / np.where(Isd > 0, Isd, 1.0), 0.0)
        m._replay(s, coef, (s * D + det).astype(np.int32), J2[0], J2[1])
    J2 = J2.permute(1, 2, 0)
    a, b = Jall.cpu().numpy(), J2.cpu().numpy()
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    m.close()


def test_several_sources_in_one_run_equal_separate_runs():
    """umcx_session_set_sources: photon p % N of source p // N is the photon a run of that source
    alone traces (same rows bitwise, same final weights)."""
    import ctypes as C
    from umcx_amd._native import check, lib
    g = O.golden("c4src")
    cfg = copy.deepcopy(g["cfg"])
    cfg["Session"]["Photons"] = 30000
    srcs = J.grid_sources(64, 64, 2, 16.5)
    N = 30000
    say("several sources in one run")
    table = np.array([list(p) + list(d) for p, d in srcs], dtype=np.float32).reshape(-1)
    with U.Simulation(cfg, tally_mode=0) as sim:
        check(lib().umcx_session_set_sources(sim._h, table.ctypes.data, len(srcs), N))
        sim.run(0, len(srcs) * N)
        rows, slots, w = sim.detp()
        check(lib().umcx_session_set_sources(sim._h, None, 0, 0))
        for s, (pos, dirv) in enumerate(srcs):
            check(lib().umcx_session_set_source(sim._h, (C.c_float * 3)(*pos), (C.c_float * 3)(*dirv)))
            sim.reset()
            sim.run(0, N)
            r1, i1, w1 = sim.detp()
            m = (slots // N) == s
            got = np.column_stack([rows[m], (slots[m] - s * N).astype(np.float32), w[m]])
            ref = np.column_stack([r1, i1.astype(np.float32), w1])
            assert np.array_equal(srt(got).view(np.uint32), srt(ref).view(np.uint32)), s


def test_run_past_the_last_source_is_rejected_before_launch():
    """With a source table set, a photon range that ends past nsrc x photons_per_source is an
    error returned by umcx_session_run on the host; no kernel is launched for it."""
    from umcx_amd._native import check, lib
    g = O.golden("c4src")
    cfg = copy.deepcopy(g["cfg"])
    srcs = J.grid_sources(64, 64, 2, 16.5)
    N = 1000
    say("source table range check")
    table = np.array([list(p) + list(d) for p, d in srcs], dtype=np.float32).reshape(-1)
    with U.Simulation(cfg, tally_mode=0) as sim:
        check(lib().umcx_session_set_sources(sim._h, table.ctypes.data, len(srcs), N))
        total = len(srcs) * N
        for first, count in ((0, total + 1), (total, 1), (total - 5, 6), (2 ** 63, 2 ** 63)):
            rc = lib().umcx_session_run(sim._h, first, count, None)
            assert rc != 0, (first, count)
            assert b"exceed" in lib().umcx_last_error(), (first, count)
        # nothing ran: tallies are still zero
        assert sim.energies()[1] == 0.0
        check(lib().umcx_session_set_sources(sim._h, None, 0, 0))


@pytest.mark.parametrize("nside,photons", [(1, 200000), (2, 100000)])
def test_jacobian_equals_oracle_replay_entry_by_entry(nside, photons):
    """Every entry of the GPU Jacobian (1 or 4 sources x 16 detectors x 131072 voxels x {mua,
    mus}) against the oracle's per-voxel replay of the same detected photons with the same float32
    coefficients (w / I_sd), accumulated in float64. The oracle also sums the magnitudes each
    entry is formed from (|coef| L for mua, |coef| (n/mus + L) for mus); float32 rounding of the
    terms and of the atomic sums stays below 2e-5 of that scale, and entries no segment reaches
    are exactly zero on both sides."""
    g = O.golden("c4src")
    srcs = J.grid_sources(64, 64, nside, 16.5)
    say(f"jacobian vs oracle replay, {len(srcs)} sources x {photons} photons")
    m = J.MultiSourceModel(g["cfg"], srcs, photons)
    m.forward()
    Jg = m.jacobian().cpu().numpy()                     # (S*D, nvox, 2)
    S, D = len(srcs), m.ndet
    ids, sidx, sl, cf = [], [], [], []
    for s, (i, det, w) in enumerate(m.records):
        Isd = m.I[s, det]
        coef = np.where(Isd > 0, w / np.where(Isd > 0, Isd, 1.0), 0.0).astype(np.float32)
        ids.append(i.astype(np.uint32)); sidx.append(np.full(len(i), s, np.int32))
        sl.append((s * D + det).astype(np.int32)); cf.append(coef.astype(np.float64))
    m.close()
    lab = U.domain(m.cfg)
    prob = O.problem_from_config(m.cfg, lab)
    table = np.array([list(p) + list(d) for p, d in srcs], dtype=np.float32)
    r = O.replay(prob, np.concatenate(ids), np.concatenate(sl), np.concatenate(cf), S * D, sources=table,
                 src_index=np.concatenate(sidx), nthreads=8)
    nrec = sum(len(x) for x in ids)
    assert nrec > 1000
    for k, (ora, scale) in enumerate(((r["gmua"], r["amua"]), (r["gmus"], r["amus"]))):
        gpu = Jg[:, :, k].astype(np.float64)
        untouched = scale == 0
        assert np.all(gpu[untouched] == 0.0), k
        ratio = np.abs(gpu[~untouched] - ora[~untouched]) / scale[~untouched]
        say(f"  param {k}: {int((~untouched).sum())} entries, max |gpu - oracle| / scale = {ratio.max():.3e}")
        assert ratio.max() <= 2e-5, (k, ratio.max())
