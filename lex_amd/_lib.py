"""ctypes binding of liblexhip.so (the C ABI in include/lex_capi.h).

The product path runs only through this library: if it is missing or a call fails,
an exception is raised.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblexhip.so")
CSRC = os.path.join(_HERE, "csrc")
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


class LexError(RuntimeError):
    """A liblexhip call returned a non-zero status."""


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float)]


class StepOpts(C.Structure):
    _fields_ = [("dt", C.c_float), ("n_sound", C.c_int), ("turb", C.c_int), ("cs", C.c_float),
                ("prandtl", C.c_float), ("n_passive", C.c_int), ("impl", C.c_int), ("solver", C.c_int),
                ("damp", C.c_int), ("rad", C.c_int), ("z_damping", C.c_float), ("rd_alpha", C.c_float),
                ("sfc", C.c_int), ("cor", C.c_int), ("fcor", C.c_float), ("u_geo", C.c_float), ("v_geo", C.c_float)]


# int fn(int stage, float* const* stage_state, float* sgs_u, float* sgs_v, float* sgs_w,
#        float* const* sgs_scalars, void* user): the per-stage SGS function of
# lex_rk3_step_stage_sgs
STAGE_SGS_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p,
                           C.POINTER(C.c_void_p), C.c_void_p)


class BaseCols(C.Structure):
    _fields_ = [("rho0", C.c_void_p), ("theta0", C.c_void_p), ("qv0", C.c_void_p), ("pi0", C.c_void_p),
                ("rho0_theta0", C.c_void_p), ("surface_t", C.c_void_p)]


def build(verbose: bool = False) -> str:
    """Compile liblexhip.so for gfx950 in-tree (one hipcc invocation, unity build)."""
    src = os.path.join(CSRC, "lex_capi.hip")
    out = LIB_PATH + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, "-o", out, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise LexError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(out, LIB_PATH)
    if verbose:
        print(" ".join(cmd))
    return LIB_PATH


_lib = None

_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
_G = C.POINTER(Grid)

# name -> argtypes (all return int unless listed in _RESTYPE)
_SIGS = {
    "lex_last_error": [],
    "lex_abi_version": [],
    "lex_pad": [_I, _I, _I, _I, _I, _P, _P, _P],
    "lex_set_w_bc": [_G, _P, _P],
    "lex_div_rho": [_G, _P, _P, _P, _P, _P, _P],
    "lex_div": [_G, _P, _P, _P, _P, _P],
    "lex_adv_scalar": [_G, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P],
    "lex_adv_u": [_G, _P, _P, _P, _P, _F, _P, _P, _P, _P],
    "lex_adv_v": [_G, _P, _P, _P, _P, _F, _P, _P, _P, _P],
    "lex_adv_w": [_G, _P, _P, _P, _P, _F, _P, _P, _P],
    "lex_adv_pi": [_G, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_buoyancy": [_G, _P, _P, _P, _P, _P, _P, _P],
    "lex_theta_rho": [C.c_long, _P, _P, _P, _P],
    "lex_pgf": [_G, _P, _P, _P, _P, _P, _P],
    "lex_coriolis": [_G, _P, _P, _F, _P, _P, _P],
    "lex_coriolis_lspgrad": [_G, _P, _P, _F, _F, _F, _P, _P, _P],
    "lex_momentum_rhs": [_G] + [_P] * 12 + [_P],
    "lex_rtt": [_G, _P, _P, _P, _P, _P, _P],
    "lex_p_rhs": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_p_matvec": [_G, _P, _P, _P, _P],
    "lex_bicgstab_workspace_bytes": [_G],
    "lex_bicgstab": [_G, _P, _P, _P, _P, _D, _D, _I, _P, C.c_size_t, C.POINTER(_I), _P],
    "lex_bicgstab_noalloc": [_G, _P, _P, _P, _P, _D, _D, _I, C.POINTER(_I), _P],
    "lex_acoustic_substep": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P],
    "lex_acoustic_substep_pair": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P],
    "lex_dl_pack": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_dl_stagger": [_G, _P, _P, _P, _P, _P, _P, _P],
    "lex_dl_apply": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_laplace": [_G, C.c_long, _P, _P, _P],
    "lex_weighted_sum_workspace_bytes": [],
    "lex_weighted_sum_grad": [C.c_long, _I, _P, _P, _P, C.c_long, _P, _P],
    "lex_dl_apply_vjp": [_G] + [_P] * 12,
    "lex_dl_pack_vjp": [_G] + [_P] * 9,
    "lex_weighted_sum": [C.c_long, _I, _P, _P, _P, C.c_long, _P, _P, C.c_size_t, _P],
    "lex_atm_ocn_flux": [C.c_long, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_surface_fluxes": [_G, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_scalar_update": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _F, _P, _P, _P],
    "lex_momentum_update": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P, _P, _P],
    "lex_smag_deformation": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_smag_nm": [_G, _P, _P, _P, _P],
    "lex_smag_k": [_G, _P, _P, _F, _F, _P, _P, _P],
    "lex_smag_momentum": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_smag_scalar": [_G, _P, _P, _P, _P, _P],
    "lex_smag_k2": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P],
    "lex_smag2_momentum": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_smag2_scalar": [_G, _P, _P, _P, _P, _P, _P],
    "lex_base_state": [_I, _D, _D, _P, _P, _I, _D, C.POINTER(_I), _P],
    "lex_rk3_workspace_bytes": [_G, _I],
    "lex_rk3_step": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), _P, C.c_size_t, _P],
    "lex_profile_enable": [_I, _I],
    "lex_profile_read": [C.POINTER(_D), C.POINTER(_I), _I],
    "lex_profile_kernel_name": [_I],
    # adjoint (vector-Jacobian products); every *_vjp accumulates into its outputs
    "lex_fold": [_G, _I, _P, _P],
    "lex_adv_scalar_vjp": [_G, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P],
    "lex_adv_u_vjp": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_adv_v_vjp": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_adv_w_vjp": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_buoyancy_vjp": [_G, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_theta_rho_vjp": [C.c_long, _P, _P, _P, _P, _P, _P],
    "lex_acoustic_substep_vjp": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F] + [_P] * 16 + [_P],
    "lex_smag_vjp_workspace_bytes": [_G],
    "lex_smag_vjp": [_G, _P, _P, _P, _P, _P, _P, _I, C.POINTER(_P), _F, _F, _P, _P, _P, C.POINTER(_P),
                     _P, _P, _P, _P, _P, C.POINTER(_P), _P, _P, C.c_size_t, _P],
    "lex_rk3_vjp_workspace_bytes": [_G, _F, _I, _I],
    "lex_rk3_step_vjp": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), C.POINTER(_P),
                         C.POINTER(_P), _P, _P, C.c_size_t, _P],
    "lex_rk3_tape_bytes": [_G, _F, _I, _I],
    "lex_rk3_adjoint_bytes": [_G, _F, _I, _I],
    "lex_rk3_step_tape": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), C.POINTER(_P), _P,
                          C.c_size_t, _P],
    "lex_rk3_step_vjp_taped": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), _P, C.c_size_t,
                               C.POINTER(_P), C.POINTER(_P), _P, _P, C.c_size_t, _P],
    # learned SGS net: BF16 MFMA convolutions
    "lex_conv3d_packed_weight_bytes": [_I, _I, _I],
    "lex_conv3d_pack_weights": [_P, _I, _I, _I, _P, _P],
    "lex_conv3d_bf16_fwd": [_I, _I, _I, _I, _I, _P, _P, _P, _I, _I, _P, _P, _P, _P],
    "lex_conv3d_bf16_fwd_circular": [_I, _I, _I, _I, _I, _P, _P, _P, _I, _I, _P, _P, _P, _P],
    "lex_conv3d_bf16_bwd_data": [_I, _I, _I, _I, _I, _P, _P, _P, _P, _P],
    "lex_conv3d_bf16_bwd_data_gelu": [_I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P],
    "lex_conv3d_bf16_bwd_weight": [_I, _I, _I, _I, _I, _P, _P, _P, _P, _P],
    "lex_conv3d_f32": [_P, _I, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _I, _I, _I, _I, _P],
    "lex_gelu_bwd": [C.c_long, _I, _I, _I, _P, _I, _P, _I, _P, _P],
    "lex_to_bf16_channels_last": [C.c_long, _I, _I, _I, _P, _P, _P],
    "lex_sgs_features": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_sgs_apply": [_G, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_coarsen": [_G, _I, _I, _I, _I, _P, _P, _P],
    "lex_rk3_step_ext_sgs": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), _P, _P, _P,
                             C.POINTER(_P), _P, C.c_size_t, _P],
    "lex_rk3_workspace_sgs_offsets": [_G, _I, C.POINTER(C.c_size_t)],
    "lex_rk3_step_stage_sgs": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), STAGE_SGS_FN, _P, _P,
                               C.c_size_t, _P],
    "lex_acoustic_substep_implicit": [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _F, _P, _P, _P, _P, _P],
    # training the net through the solver
    "lex_sgs_apply_vjp": [_G, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_sgs_features_vjp": [_G, _P, _P, _P, _P, _P, _P, _P, _P],
    "lex_rk3_step_tape_ext_sgs": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), _P, _P, _P,
                                  C.POINTER(_P), C.POINTER(_P), _P, C.c_size_t, _P],
    "lex_rk3_step_vjp_taped_ext_sgs": [_G, C.POINTER(StepOpts), C.POINTER(BaseCols), C.POINTER(_P), _P,
                                       C.c_size_t, C.POINTER(_P), C.POINTER(_P), _P, _P, _P, C.POINTER(_P), _P,
                                       C.c_size_t, _P],
}
_RESTYPE = {"lex_last_error": C.c_char_p, "lex_bicgstab_workspace_bytes": C.c_size_t,
            "lex_rk3_workspace_bytes": C.c_size_t, "lex_profile_kernel_name": C.c_char_p,
            "lex_smag_vjp_workspace_bytes": C.c_size_t, "lex_weighted_sum_workspace_bytes": C.c_size_t, "lex_rk3_vjp_workspace_bytes": C.c_size_t,
            "lex_conv3d_packed_weight_bytes": C.c_size_t, "lex_rk3_tape_bytes": C.c_size_t,
            "lex_rk3_adjoint_bytes": C.c_size_t}


def profile_enable(on: bool, capacity_pairs: int = 8192):
    call("lex_profile_enable", int(on), int(capacity_pairs))


def profile_read(n: int = 32):
    """{kernel name: (total ms, launches)} since the last read (synchronises)."""
    lib = load()
    ms = (C.c_double * n)()
    cnt = (C.c_int * n)()
    k = lib.lex_profile_read(ms, cnt, n)
    if k < 0:
        raise LexError("lex_profile_read failed: " + lib.lex_last_error().decode(errors="replace"))
    out = {}
    for i in range(min(k, n)):
        if cnt[i]:
            out[lib.lex_profile_kernel_name(i).decode()] = (ms[i], cnt[i])
    return out


def load():
    """Load liblexhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LexError(f"{LIB_PATH} is missing: build it with lex_amd._lib.build() "
                       "(or __graft_entry__.build()); there is no fallback path")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


# tensors whose pointers were taken for the call being assembled: a temporary (say
# `ptr(x.contiguous())`) stays alive until the call has been issued, so a later allocation
# in the same argument list cannot take its memory first; after the call, reuse of the
# memory is ordered on the stream behind the launched kernels
_KEEP: list = []
_DEPTH = [0]


def call(name, *args):
    """Call an entry point; raise LexError with lex_last_error() on a non-zero status."""
    lib = load()
    _DEPTH[0] += 1
    try:
        rc = getattr(lib, name)(*args)
    finally:
        _DEPTH[0] -= 1
        if _DEPTH[0] == 0:   # calls made from inside a callback keep the outer call's tensors
            _KEEP.clear()
    if rc != 0:
        msg = lib.lex_last_error().decode(errors="replace")
        raise LexError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t: torch.Tensor | None):
    """Device pointer of a contiguous CUDA/HIP float tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise LexError("liblexhip needs device tensors (got a CPU tensor)")
    if not t.is_contiguous():
        raise LexError("liblexhip needs contiguous tensors")
    _KEEP.append(t)
    return t.data_ptr()


def stream():
    return torch.cuda.current_stream().cuda_stream


def grid(nx, ny, nz, dx, dy, dz) -> Grid:
    return Grid(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz))
