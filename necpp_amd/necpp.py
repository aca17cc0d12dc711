"""Python binding of libnecpp_amd.so with the reference's C API names.

``from necpp_amd.necpp import *`` gives the same functions the external
``necpp`` module exposes to nec2++'s Python tests (nec_create, nec_wire, ...,
nec_impedance_real), plus the GPU extras of include/necpp_gpu.h.  Every call
goes through the C ABI; there is no Python or CPU fallback: loading fails
loudly when the library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnecpp_amd.so")

_lib = None

# (name, restype, argtypes) of every exported entry point
_D, _I, _L, _P, _LL, _U = ctypes.c_double, ctypes.c_int, ctypes.c_long, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_uint
_SIGS = [
    ("nec_create", _P, []),
    ("nec_delete", _L, [_P]),
    ("nec_benchmark", _L, []),
    ("nec_wire", _L, [_P, _I, _I] + [_D] * 9),
    ("nec_sp_card", _L, [_P, _I] + [_D] * 6),
    ("nec_sc_card", _L, [_P, _I] + [_D] * 6),
    ("nec_gm_card", _L, [_P, _I, _I] + [_D] * 6 + [_I]),
    ("nec_gx_card", _L, [_P, _I, _I]),
    ("nec_geometry_complete", _L, [_P, _I]),
    ("nec_error_message", ctypes.c_char_p, []),
    ("nec_medium_parameters", _L, [_P, _D, _D]),
    ("nec_gn_card", _L, [_P, _I, _I] + [_D] * 6),
    ("nec_fr_card", _L, [_P, _I, _I, _D, _D]),
    ("nec_ek_card", _L, [_P, _I]),
    ("nec_ld_card", _L, [_P, _I, _I, _I, _I, _D, _D, _D]),
    ("nec_ex_card", _L, [_P, _I, _I, _I, _I] + [_D] * 6),
    ("nec_excitation_voltage", _L, [_P, _I, _I, _D, _D]),
    ("nec_excitation_current", _L, [_P] + [_D] * 6),
    ("nec_excitation_planewave", _L, [_P, _I, _I] + [_D] * 6),
    ("nec_tl_card", _L, [_P, _I, _I, _I, _I] + [_D] * 6),
    ("nec_nt_card", _L, [_P, _I, _I, _I, _I] + [_D] * 6),
    ("nec_xq_card", _L, [_P, _I]),
    ("nec_gd_card", _L, [_P, _D, _D, _D, _D]),
    ("nec_rp_card", _L, [_P, _I, _I, _I, _I, _I, _I, _I] + [_D] * 6),
    ("nec_pt_card", _L, [_P, _I, _I, _I, _I]),
    ("nec_pq_card", _L, [_P, _I, _I, _I, _I]),
    ("nec_kh_card", _L, [_P, _D]),
    ("nec_ne_card", _L, [_P, _I, _I, _I, _I] + [_D] * 6),
    ("nec_nh_card", _L, [_P, _I, _I, _I, _I] + [_D] * 6),
    ("nec_cp_card", _L, [_P, _I, _I, _I, _I]),
    ("nec_pl_card", _L, [_P, ctypes.c_char_p, _I, _I, _I, _I]),
    ("nec_gain", _D, [_P, _I, _I, _I]),
] + [(f"nec_gain{s}_{k}", _D, [_P, _I]) for s in ("", "_rhcp", "_lhcp") for k in ("max", "min", "mean", "sd")] + [
    ("nec_impedance_real", _D, [_P, _I]),
    ("nec_impedance_imag", _D, [_P, _I]),
    ("nec_impedance_real_indexed", _D, [_P, _I, _I]),
    ("nec_impedance_imag_indexed", _D, [_P, _I, _I]),
    ("nec_gpu_info", _L, [ctypes.c_char_p, _I]),
    ("nec_gpu_segment_count", _L, [_P]),
    ("nec_gpu_segments", _L, [_P, _P, _P, _L]),
    ("nec_gpu_set_precision", _L, [_P, _I]),
    ("nec_gpu_batch_solve", _L, [_P, _P, _P, _P]),
    ("nec_gpu_batch_adjoint", _L, [_P, _D, _D, _D, _D, _D, _P, _P, _P, _P, _P]),
    ("nec_gpu_set_refine_tolerance", _L, [_P, _D]),
    ("nec_gpu_refinement", _L, [_P, _P, _P]),
    ("nec_gpu_factor_fallbacks", _L, [_P, _P, _P]),
    ("nec_gpu_adjoint_gradient", _L, [_P, _D, _D, _D, _D, _D, _P, _P, _P, _P]),
    ("nec_gpu_currents", _L, [_P, _I, _P, _L]),
    ("nec_gpu_stage_times", _L, [_P, _P]),
    ("nec_gpu_benchmark_solve", _L, [_P, _I, _I, _P]),
    ("nec_gpu_kernel_times", _L, [_P, _P, _L]),
    ("nec_gpu_set_kernel_timing", _L, [_P, _I]),
    ("nec_gpu_kernel_name", ctypes.c_char_p, [_I]),
    ("nec_gpu_fill_matrix", _L, [_P, _D, _P, _L]),
    ("nec_gpu_pattern_array", _L, [_P, _I, _I, _P, _L]),
    ("nec_gpu_pattern_average", _L, [_P, _I, _P, _P]),
    ("necgpu_zgetrf_workspace", _LL, [_I]),
    ("necgpu_zgetrf", _L, [_P, _LL, _I, _P, _P, _P, _P]),
    ("necgpu_cgetrf", _L, [_P, _LL, _I, _P, _P, _P, _I, _P]),
    ("necgpu_zgetrf_noalloc", _L, [_P, _LL, _I, _P, _P, _P]),
    ("nec_gpu_deck", _L, [_P, ctypes.c_char_p]),
    ("nec_gpu_near_field", _L, [_P, _I, _P, _L]),
    ("nec_gpu_near_field_kind", _L, [_P, _I]),
    ("necgpu_check_mfma_f64", _L, [_P, _P, _P, _P]),
    ("necgpu_check_mfma_f32", _L, [_P, _P, _P, _P]),
    ("necgpu_zgetrs_workspace", _LL, [_I]),
    ("necgpu_zgetrs", _L, [_P, _LL, _I, _P, _P, _P, _U, _P]),
]

EXPORTS = [s[0] for s in _SIGS]


def lib():
    """Load libnecpp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m necpp_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _wrap(name):
    def f(*args):
        return getattr(lib(), name)(*args)
    f.__name__ = name
    return f


for _name in EXPORTS:
    if _name not in ("nec_error_message",):
        globals()[_name] = _wrap(_name)


def nec_error_message():
    m = lib().nec_error_message()
    return m.decode() if m else ""


# ------------------------------------------------------------------ helpers

def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def gpu_currents(ctx, freq_index=0):
    n = lib().nec_gpu_segment_count(ctx)
    out = np.zeros(2 * n)
    if lib().nec_gpu_currents(ctx, freq_index, _ptr(out), n) != 0:
        raise RuntimeError(nec_error_message())
    return out[0::2] + 1j * out[1::2]


def gpu_segments(ctx):
    """(table n x 8, icon1, icon2, tags) of a completed geometry."""
    n = lib().nec_gpu_segment_count(ctx)
    out = np.zeros(8 * n)
    conn = np.zeros(3 * n, np.int32)
    if lib().nec_gpu_segments(ctx, _ptr(out), _ptr(conn), n) != 0:
        raise RuntimeError(nec_error_message())
    return out.reshape(n, 8), conn[:n].copy(), conn[n:2 * n].copy(), conn[2 * n:].copy()


def gpu_stage_times(ctx):
    out = np.zeros(4)
    if lib().nec_gpu_stage_times(ctx, _ptr(out)) != 0:
        raise RuntimeError(nec_error_message())
    return dict(zip(("fill_ms", "factor_ms", "solve_ms", "pattern_ms"), out.tolist()))


def gpu_fill_matrix(ctx, freq_mhz):
    n = lib().nec_gpu_segment_count(ctx)
    out = np.zeros(2 * n * n)
    if lib().nec_gpu_fill_matrix(ctx, freq_mhz, _ptr(out), n) != 0:
        raise RuntimeError(nec_error_message())
    return (out[0::2] + 1j * out[1::2]).reshape(n, n)


PATTERN_ARRAYS = ("gain", "pg_vert", "pg_horiz", "pg_tot", "pg_rhcp", "pg_lhcp", "axial_ratio", "tilt", "sense",
                  "avg_scale", "e_theta", "e_phi")


def gpu_pattern_array(ctx, freq_index, name, npoints):
    which = PATTERN_ARRAYS.index(name)
    cplx = name in ("e_theta", "e_phi")
    out = np.zeros(npoints * (2 if cplx else 1))
    if lib().nec_gpu_pattern_array(ctx, freq_index, which, _ptr(out), out.size) != 0:
        raise RuntimeError(nec_error_message())
    return out[0::2] + 1j * out[1::2] if cplx else out


def gpu_kernel_times(ctx):
    """{kernel family: (launches, total ms)} of the last nec_gpu_benchmark_solve's timed steps."""
    out = np.zeros(64)
    k = lib().nec_gpu_kernel_times(ctx, _ptr(out), out.size)
    if k < 0:
        raise RuntimeError(nec_error_message())
    return {lib().nec_gpu_kernel_name(i).decode(): (int(out[2 * i]), float(out[2 * i + 1])) for i in range(k)}


def gpu_refinement(ctx):
    """(steps, relative residual) of the last mixed-precision solve."""
    steps = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    if lib().nec_gpu_refinement(ctx, ctypes.byref(steps), ctypes.byref(res)) != 0:
        raise RuntimeError(nec_error_message())
    return steps.value, res.value


def gpu_factor_fallbacks(ctx):
    """(lu, pivot): factorisations repeated after a timed-out panel
    exchange, and mixed precision's fallbacks to partial pivoting."""
    lu = ctypes.c_long(0)
    pv = ctypes.c_long(0)
    if lib().nec_gpu_factor_fallbacks(ctx, ctypes.byref(lu), ctypes.byref(pv)) != 0:
        raise RuntimeError(nec_error_message())
    return lu.value, pv.value


ADJOINT_INFO = ("gain_db", "zin_re", "zin_im", "host_ms", "device_ms", "dJ_dlambda")


def gpu_adjoint_gradient(ctx, theta_deg=90.0, phi_deg=0.0, weight=0.01, z0=50.0):
    """Adjoint gradient of J = -G(theta, phi) + weight |Z_in - z0|^2 for the
    last solve of ctx.  Returns (J, grad [n, 7] per metre: x1 y1 z1 x2 y2 z2
    radius of every segment, dJ/d(frequency MHz), info dict)."""
    n = lib().nec_gpu_segment_count(ctx)
    J = ctypes.c_double(0.0)
    df = ctypes.c_double(0.0)
    g = np.zeros(7 * max(n, 1))
    info = np.zeros(6)
    z0 = complex(z0)
    if lib().nec_gpu_adjoint_gradient(ctx, theta_deg, phi_deg, weight, z0.real, z0.imag, ctypes.byref(J), _ptr(g),
                                      ctypes.byref(df), _ptr(info)) != 0:
        raise RuntimeError(nec_error_message())
    return J.value, g[:7 * n].reshape(n, 7), df.value, dict(zip(ADJOINT_INFO, info.tolist()))


class BatchDesc(ctypes.Structure):
    """struct nec_gpu_batch of necpp_gpu.h."""
    _fields_ = [("n_designs", ctypes.c_int), ("wire_ptr", _P), ("wire_tag", _P), ("wire_ns", _P),
                ("wire_geom", _P), ("src_tag", _P), ("src_seg", _P), ("src_volt", _P), ("n_freq", ctypes.c_int),
                ("freq_mhz", _P), ("n_dir", ctypes.c_int), ("theta_deg", _P), ("phi_deg", _P),
                ("rkh", ctypes.c_double), ("panel_width", ctypes.c_int), ("repeats", ctypes.c_int)]


BATCH_STATS = ("solves", "segments", "host_ms", "fill_ms", "lu_ms", "solve_pattern_ms", "graph_ms_per_replay",
               "panel_width", "autotune_ms_p8", "autotune_ms_p16", "autotune_ms_p32", "graph_matches_eager")


def _batch_desc(decks, freqs, theta_deg=(), phi_deg=(), repeats=0, panel_width=0, rkh=1.0):
    """nec_gpu_batch of necpp_amd.decks structures (and the arrays it points
    to, which must stay alive while it is used)."""
    nd = len(decks)
    freqs = np.asarray(freqs, dtype=float)
    if freqs.ndim == 1:
        freqs = np.tile(freqs, (nd, 1))
    nf = freqs.shape[1]
    wire_ptr, tags, ns, geom = [0], [], [], []
    for d in decks:
        for w in d["wires"]:
            tags.append(w[0])
            ns.append(w[1])
            geom.extend(w[2:9])
        wire_ptr.append(len(tags))
    a = {"wire_ptr": np.asarray(wire_ptr, np.int32), "wire_tag": np.asarray(tags, np.int32),
         "wire_ns": np.asarray(ns, np.int32), "wire_geom": np.asarray(geom, float),
         "src_tag": np.asarray([d["source"][0] for d in decks], np.int32),
         "src_seg": np.asarray([d["source"][1] for d in decks], np.int32),
         "src_volt": np.asarray([[complex(d["voltage"]).real, complex(d["voltage"]).imag] for d in decks], float),
         "freq_mhz": np.ascontiguousarray(freqs), "theta_deg": np.asarray(theta_deg, float),
         "phi_deg": np.asarray(phi_deg, float)}
    ndir = len(a["theta_deg"])
    b = BatchDesc(nd, _ptr(a["wire_ptr"]), _ptr(a["wire_tag"]), _ptr(a["wire_ns"]), _ptr(a["wire_geom"]),
                  _ptr(a["src_tag"]), _ptr(a["src_seg"]), _ptr(a["src_volt"]), nf, _ptr(a["freq_mhz"]), ndir,
                  _ptr(a["theta_deg"]), _ptr(a["phi_deg"]), rkh, panel_width, repeats)
    return b, a, nd, nf, ndir


BATCH_ADJOINT_STATS = ("solves", "segments", "host_prep_ms", "fill_ms", "lu_ms", "solve_ms", "host_objective_ms",
                       "device_adjoint_ms", "host_gradient_ms", "total_ms")


def gpu_batch_adjoint(decks, freqs, theta_deg=90.0, phi_deg=0.0, weight=0.01, z0=50.0, panel_width=0, rkh=1.0):
    """nec_gpu_batch_adjoint over necpp_amd.decks structures.  Returns (J
    [design, freq], grads: list over solves of [n, 7] arrays, dJ/dfreq
    [design, freq], info [design, freq, 3], stats)."""
    b, arrs, nd, nf, _ = _batch_desc(decks, freqs, panel_width=panel_width, rkh=rkh)
    nseg = [sum(w[1] for w in d["wires"]) for d in decks]
    total = sum(nseg) * nf
    J = np.zeros(nd * nf)
    grad = np.zeros(7 * max(total, 1))
    df = np.zeros(nd * nf)
    info = np.zeros(3 * nd * nf)
    stats = np.zeros(10)
    z0 = complex(z0)
    if lib().nec_gpu_batch_adjoint(ctypes.byref(b), theta_deg, phi_deg, weight, z0.real, z0.imag, _ptr(J), _ptr(grad),
                                   _ptr(df), _ptr(info), _ptr(stats)) != 0:
        raise RuntimeError(nec_error_message())
    grads, off = [], 0
    for d in range(nd):
        for _ in range(nf):
            grads.append(grad[7 * off:7 * (off + nseg[d])].reshape(nseg[d], 7))
            off += nseg[d]
    return (J.reshape(nd, nf), grads, df.reshape(nd, nf), info.reshape(nd, nf, 3),
            dict(zip(BATCH_ADJOINT_STATS, stats.tolist())))


def gpu_batch_solve(decks, freqs, theta_deg=(), phi_deg=(), repeats=0, panel_width=0, rkh=1.0):
    """Batched solves of necpp_amd.decks structures (GW wires, one EX 0
    source each).  freqs: one list for all designs, or one row per design.
    Returns (zin complex [design, freq], gain [design, freq, dir], stats)."""
    b, arrs, nd, nf, ndir = _batch_desc(decks, freqs, theta_deg, phi_deg, repeats, panel_width, rkh)
    zin = np.zeros(2 * nd * nf)
    gain = np.zeros(max(1, nd * nf * ndir))
    stats = np.zeros(12)
    if lib().nec_gpu_batch_solve(ctypes.byref(b), _ptr(zin), _ptr(gain) if ndir else None, _ptr(stats)) != 0:
        raise RuntimeError(nec_error_message())
    z = (zin[0::2] + 1j * zin[1::2]).reshape(nd, nf)
    return z, gain[:nd * nf * ndir].reshape(nd, nf, ndir), dict(zip(BATCH_STATS, stats.tolist()))


def gpu_info():
    buf = ctypes.create_string_buffer(256)
    lib().nec_gpu_info(buf, 256)
    return buf.value.decode()
