"""Differentiable Yagi-Uda design objective over the C ABI.

J(p) = -G(theta=90, phi=0) + weight * |Z_in - z0|^2 for the parameter rows
of decks.c4_params (half-lengths, spacings, radii in wavelengths).  Each
evaluation is one FP64 solve through the libnecpp C API (fill, LU, solve on
the GPU) followed by nec_gpu_adjoint_gradient, which reuses the LU factors
for the transposed solve and contracts lam^T (dZ/dp) x on the GPU.  The
segment-endpoint gradient it returns is mapped onto the design parameters
here (the endpoints are linear in them).  YagiObjective exposes this as a
torch.autograd.Function.
"""
import time

import numpy as np

from . import decks
from . import necpp as api


def params_gradient(p, grad7, freq_mhz=decks.ADJOINT_FREQ_MHZ, seg_per_element=31):
    """dJ/dp from the per-segment endpoint gradient (nec_gpu_adjoint_gradient)
    for the deck decks.c4_deck(p): element e's segment k runs from
    y = -h_e + 2 k h_e / ns to y = -h_e + 2 (k + 1) h_e / ns at
    x = sum of the spacings before e, with radius r_e."""
    lam = decks.c4_lambda(freq_mhz)
    ne = (len(p) + 1) // 3
    g = np.asarray(grad7).reshape(ne, seg_per_element, 7)
    k = np.arange(seg_per_element)
    w1 = -1.0 + 2.0 * k / seg_per_element
    w2 = -1.0 + 2.0 * (k + 1) / seg_per_element
    dh = lam * (g[:, :, 1] @ w1 + g[:, :, 4] @ w2)
    dx = g[:, :, 0].sum(axis=1) + g[:, :, 3].sum(axis=1)
    # spacing i moves every element after it
    dd = lam * np.cumsum(dx[::-1])[::-1][1:]
    dr = lam * g[:, :, 6].sum(axis=1)
    return np.concatenate([dh, dd, dr])


def _solve(deck, ctx, with_rp):
    decks.apply_to_api(deck, api, ctx, with_rp=with_rp)


def objective(p, weight=0.01, z0=50.0, freq_mhz=decks.ADJOINT_FREQ_MHZ, seg_per_element=31):
    """J(p) from the reference API's own outputs (RP card gain, impedance):
    the quantity the finite differences act on."""
    deck = decks.c4_deck(p, freq_mhz, seg_per_element)
    ctx = api.nec_create()
    try:
        _solve(deck, ctx, True)
        g = api.nec_gain(ctx, 0, 0, 0)
        z = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
    finally:
        api.nec_delete(ctx)
    return -g + weight * abs(z - complex(z0)) ** 2


def objective_and_gradient(p, theta_deg=90.0, phi_deg=0.0, weight=0.01, z0=50.0, freq_mhz=decks.ADJOINT_FREQ_MHZ,
                           seg_per_element=31):
    """(J, dJ/dp, info) of one design by the adjoint method."""
    deck = decks.c4_deck(p, freq_mhz, seg_per_element)
    ctx = api.nec_create()
    try:
        t0 = time.perf_counter()
        _solve(deck, ctx, False)
        t1 = time.perf_counter()
        J, g7, dfreq, info = api.gpu_adjoint_gradient(ctx, theta_deg, phi_deg, weight, z0)
        t2 = time.perf_counter()
        info["solve_wall_ms"] = 1e3 * (t1 - t0)
        info["adjoint_wall_ms"] = 1e3 * (t2 - t1)
        info["dJ_dfreq_mhz"] = dfreq
        info.update(api.gpu_stage_times(ctx))
    finally:
        api.nec_delete(ctx)
    return J, params_gradient(p, g7, freq_mhz, seg_per_element), info


def batch_objective_and_gradient(P, theta_deg=90.0, phi_deg=0.0, weight=0.01, z0=50.0,
                                 freq_mhz=decks.ADJOINT_FREQ_MHZ, seg_per_element=31):
    """(J [B], dJ/dp [B, 44], stats) of a batch of parameter rows in one call
    of nec_gpu_batch_adjoint: one batched solve and one batched adjoint on
    the GPU for all rows."""
    P = np.atleast_2d(np.asarray(P, dtype=float))
    ds = [decks.c4_deck(p, freq_mhz, seg_per_element) for p in P]
    J, grads, _, _, stats = api.gpu_batch_adjoint(ds, [freq_mhz], theta_deg, phi_deg, weight, z0)
    G = np.array([params_gradient(p, g7, freq_mhz, seg_per_element) for p, g7 in zip(P, grads)])
    return J[:, 0], G, stats


def central_differences(p, rel_step=1e-6, per_component=False, fn=None, **kw):
    """Central finite differences of J (objective(), or fn) with the step
    rel_step * |p| (the Euclidean norm of the parameter row) for every
    component, or rel_step * |p_k| per component."""
    fn = fn or objective
    g = np.zeros(len(p))
    hn = rel_step * float(np.linalg.norm(p))
    for k in range(len(p)):
        h = rel_step * abs(p[k]) if per_component else hn
        a = np.array(p, dtype=float)
        b = a.copy()
        a[k] += h
        b[k] -= h
        g[k] = (fn(a, **kw) - fn(b, **kw)) / (a[k] - b[k])
    return g


def _torch_function():
    import torch

    class YagiObjective(torch.autograd.Function):
        """J for a batch of parameter rows [B, 3 ne - 1] (float64), all rows
        in one batched solve and adjoint; the backward pass returns the
        adjoint gradient computed with J."""

        @staticmethod
        def forward(ctx, params, weight=0.01, z0=50.0):
            P = params.detach().to("cpu", torch.float64).numpy()
            squeeze = P.ndim == 1
            J, G, _ = batch_objective_and_gradient(np.atleast_2d(P), weight=weight, z0=z0)
            Gt = torch.from_numpy(G).to(params.device, params.dtype)
            ctx.save_for_backward(Gt[0] if squeeze else Gt)
            Jt = torch.from_numpy(J).to(params.device, params.dtype)
            return Jt[0] if squeeze else Jt

        @staticmethod
        def backward(ctx, gJ):
            (G,) = ctx.saved_tensors
            if G.ndim == 1:
                return gJ * G, None, None
            return gJ[:, None] * G, None, None

    return YagiObjective


_FN = None


def yagi_objective(params, weight=0.01, z0=50.0):
    """torch entry: J of each parameter row, differentiable in params."""
    global _FN
    if _FN is None:
        _FN = _torch_function()
    return _FN.apply(params, weight, z0)
