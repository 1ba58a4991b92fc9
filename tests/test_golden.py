"""Committed golden vectors (tests/golden/rf_golden.npz, written by
tests/golden/make_golden.py): the oracle must reproduce them (CPU), and the
GPU kernels must match them through the drop-in ops (GPU)."""
import os

import numpy as np
import pytest
import torch

from oracle import lensless as L
from oracle import rf as orc
from rf_workloads import mf_scenes, tracer_scenes as ts

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "rf_golden.npz"))


def rel(x, ref):
    return float(np.linalg.norm(np.asarray(x) - ref) / np.linalg.norm(ref))


def mf_prm():
    return dict(f0=mf_scenes.MF_F0, slope=float(G["mf_as"]), fs=float(G["mf_fs"]),
                c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)


def st_prm():
    return dict(f0=ts.FC_START, slope=ts.AS_SCI, fs=float(G["mf_fs"]), c=ts.SPEED_OF_LIGHT,
                bias=ts.PATH_BIAS)


def test_oracle_reproduces_golden():
    s = G["mf_echo"]
    np.testing.assert_allclose(orc.mf_forward(s.real, s.imag, G["mf_tx"], G["mf_rx"], G["mf_vox"],
                                              **mf_prm()), G["mf_out"], rtol=1e-12, atol=0)
    K = s.shape[1]
    np.testing.assert_allclose(orc.mf_backward(G["mf_gout"], G["mf_tx"], G["mf_rx"], G["mf_vox"],
                                               K, norm=s.shape[0], **mf_prm()),
                               G["mf_grad"], rtol=1e-12, atol=0)
    out = orc.st_forward(G["st_nrm"], G["st_sig"], G["st_pts"], G["st_tx"], G["st_rx"],
                         G["st_inb"], G["st_gout"].shape[1], **st_prm())
    np.testing.assert_allclose(out, G["st_out"], rtol=1e-12, atol=0)
    on, os_, inb = orc.bilinear_forward(G["bl_n"], G["bl_s"], G["bl_z"], G["bl_tz"])
    assert np.array_equal(inb, G["bl_inb"]) and np.array_equal(on, G["bl_on"])


def _lens_batch(dev=None):
    t = lambda k: torch.tensor(G[f"ll_{k}"])  # noqa: E731
    b = ts.RayBatch(t("sdf"), t("grad"), t("dists"), t("z"), t("o"), t("d"), t("refl"),
                    t("offset"), t("n_valid"), t("inv_s"), t("power"),
                    ts.radar_for_k(int(G["ll_K"])), float(G["ll_anneal"]))
    return b if dev is None else b.to(dev)


@pytest.mark.gpu
def test_gpu_matches_golden(cuda):
    from geraf.ops.RFMatchedFilter import matched_filter_cuda
    from geraf.ops.RFSignalTracer import bilinear_raysample_cuda, signal_tracer_cuda
    from geraf.ops.RFSignalTracer.lensless_tracer import lensless_signal_tracer

    T = lambda x, **k: torch.tensor(np.asarray(x), device=cuda, **k)  # noqa: E731
    # matched filter forward + adjoint
    s = G["mf_echo"]
    A, K = s.shape
    N = G["mf_vox"].shape[0]
    re = T(s.real, requires_grad=True)
    im = T(s.imag, requires_grad=True)
    o_r, o_i = matched_filter_cuda(re, im, T(G["mf_tx"]), T(G["mf_rx"]), T(G["mf_vox"]), A, K,
                                   float(G["mf_fs"]), float(G["mf_as"]), N)
    out = o_r.detach().double().cpu().numpy() + 1j * o_i.detach().double().cpu().numpy()
    assert rel(out, G["mf_out"]) <= 1e-5
    g = G["mf_gout"]
    (o_r * T(g.real) + o_i * T(g.imag)).sum().backward()
    grad = re.grad.double().cpu().numpy() + 1j * im.grad.double().cpu().numpy()
    assert rel(grad, G["mf_grad"]) <= 1e-5
    # [A,P] tracer
    Ak, Kt = G["st_gout"].shape
    P = G["st_pts"].shape[0]
    nt = T(G["st_nrm"], requires_grad=True)
    st = T(G["st_sig"], requires_grad=True)
    o_r, o_i = signal_tracer_cuda(nt, st, T(G["st_pts"]), T(G["st_tx"]), T(G["st_rx"]),
                                  T(G["st_inb"]), Kt, float(G["mf_fs"]), ts.FC_START,
                                  ts.AS_SCI, Ak, P)
    out = o_r.detach().double().cpu().numpy() + 1j * o_i.detach().double().cpu().numpy()
    assert rel(out, G["st_out"]) <= 1e-5
    gk = G["st_gout"]
    (o_r * T(gk.real) + o_i * T(gk.imag)).sum().backward()
    assert rel(nt.grad.cpu().numpy(), G["st_gn"]) <= 1e-4
    assert rel(st.grad.cpu().numpy(), G["st_gs"]) <= 1e-4
    # bilinear
    A2, R, S = G["bl_s"].shape
    Tn = G["bl_tz"].shape[1]
    bn = T(G["bl_n"], requires_grad=True)
    bs = T(G["bl_s"], requires_grad=True)
    on, os_, inb = bilinear_raysample_cuda(bn, bs, T(G["bl_z"]), T(G["bl_tz"]), A2, R, S, Tn)
    assert np.array_equal(inb.cpu().numpy(), G["bl_inb"])
    assert np.max(np.abs(on.detach().cpu().numpy() - G["bl_on"])) < 1e-6
    assert np.max(np.abs(os_.detach().cpu().numpy() - G["bl_os"])) < 1e-6
    ((on * T(G["bl_gn"])).sum() + (os_ * T(G["bl_gs"])).sum()).backward()
    assert np.max(np.abs(bn.grad.cpu().numpy() - G["bl_rgn"])) < 1e-5
    assert np.max(np.abs(bs.grad.cpu().numpy() - G["bl_rgs"])) < 1e-5
    # fused lensless renderer
    b = _lens_batch(cuda)
    leaves = {k: getattr(b, k).clone().requires_grad_(True)
              for k in ("sdf", "grad", "refl", "inv_s", "power")}
    o_r, o_i = lensless_signal_tracer(leaves["sdf"], leaves["grad"], b.dists, b.z, b.o, b.d,
                                      leaves["inv_s"], leaves["power"],
                                      reflectance=leaves["refl"], trans_offset=b.offset,
                                      n_valid=b.n_valid, cos_anneal_ratio=b.anneal, **b.radar)
    out = o_r.detach().double().cpu().numpy() + 1j * o_i.detach().double().cpu().numpy()
    assert rel(out, G["ll_out"]) <= 1e-5
    gl = G["ll_gout"]
    (o_r * T(gl.real, dtype=torch.float32) + o_i * T(gl.imag, dtype=torch.float32)).sum() \
        .backward()
    for k in ("sdf", "grad", "refl"):
        assert rel(leaves[k].grad.double().cpu().numpy().ravel(), G[f"ll_g_{k}"].ravel()) <= 1e-4
    # inv_s / power: sums with cancellation; absolute tolerance on the scale of
    # the other gradients
    for k in ("inv_s", "power"):
        err = abs(leaves[k].grad.item() - float(G[f"ll_g_{k}"]))
        assert err <= 1e-4 * np.abs(G["ll_g_sdf"]).sum(), k
