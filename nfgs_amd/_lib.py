"""ctypes binding of libnfgs_hip.so (declarations in include/nfgs_hip.h).

The binding never falls back to anything else: if the library is missing or
fails to load, every solver call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libnfgs_hip.so"

NFGS_OK, NFGS_ERR_ARG, NFGS_ERR_HIP, NFGS_ERR_STATE, NFGS_ERR_ALLOC = 0, -1, -2, -3, -4
NFGS_F32, NFGS_F64 = 0, 1
NFGS_DCT1, NFGS_DST1 = 0, 1
NFGS_FORWARD, NFGS_INVERSE = 0, 1
(Q_SIGMA_NUM, Q_SIGMA_DEN, Q_DIV_ABS, Q_BMAG, Q_B2, Q_SRC2, Q_ZERO_MODE, Q_FACE_J2,
 Q_CELL_J2) = range(9)
NQ = 9
GD = dict(BX_MEAN=0, BY_MEAN=1, BZ_MEAN=2, D0_MEAN=3, D0_SQ=4, RECON_SQ=5, RECON_MAX=6,
          BG_INV_SQ=7, HRES_SQ=8, HOBS_SQ=9, HCONST_X=10, HCONST_Y=11, HNONCONST_SQ=12,
          HDIV_SQ=13, HCURL_SQ=14, BZ_SUM=15, BZ_ABS_SUM=16, BX_SUM=17, BY_SUM=18)
GDIAG_COUNT = 19
SD_COUNT = 10
# nfgs_field_scores sums (NFGS_SC_*)
(SC_DOT, SC_REF2, SC_SOL2, SC_COS, SC_DIFF, SC_REFN, SC_DIFF_REL, SC_JXB, SC_JN, SC_BN,
 SC_DIVABS) = range(11)
SC_COUNT = 11
# nfgs_advance_ex poloidal update modes (NFGS_POLOIDAL_*)
POLOIDAL_DIRECT, POLOIDAL_TWO_STAGE = 0, 1
# iteration diagnostics sums (NFGS_DR_*, NFGS_DF_*)
(DR_PHI_MAP, DR_PHI_REF, DR_AZ_MAP, DR_AZ_REF, DR_BCZ_RES, DR_BCZ_SRC, DR_JZ_RES, DR_JZ_SRC,
 DR_POL_RES, DR_POL_D2Z, DR_POL_LAP, DR_POL_SRC, DR_TOR_RES, DR_TOR_D2Z, DR_TOR_LAP, DR_TOR_SRC,
 DR_MIX_RES, DR_MIX_PDE, DR_MIX_DER, DR_MIX_SRC, DR_COUNT) = range(21)
DF_WEAK, DF_CLIP, DF_IMPRINT, DF_OBS, DF_COUNT = range(5)
# lateral transform arithmetic of FP32 plans (NFGS_LATERAL_*)
LATERAL_EXACT, LATERAL_BF16X3 = 0, 1
# coupled horizontal-harmonic closures (NFGS_K0_*)
K0_COUPLED, K0_SPLINE = 1, 2
K0_STAGGERED, K0_LEGACY = 0, 1
# polarity masks of the bottom parallel current (NFGS_POLARITY_*)
POLARITY = {"both": 0, "positive": 1, "negative": 2}
# learned start (NFGS_FNO_*)
NFGS_FNO_IDENTITY, NFGS_FNO_GELU = 0, 1
NFGS_FNO_PROJECT, NFGS_FNO_EXPAND = 0, 1
# nfgs_plan_set_tuning knobs (NFGS_TUNE_*)
TUNE_ZCHUNKS, TUNE_TRANSFORM_BLOCKS_PER_CU = 0, 1

vp = C.c_void_p


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dtype", C.c_int), ("rho", C.c_double), ("alpha_max", C.c_double),
                ("weak_field_floor_b2", C.c_double),
                ("win_z0", C.c_int), ("win_z1", C.c_int), ("win_y0", C.c_int),
                ("win_y1", C.c_int), ("win_x0", C.c_int), ("win_x1", C.c_int)]


class Boundary(C.Structure):
    _fields_ = [(n, vp) for n in ("b_bg", "bz_obs", "jz_t", "phi_bottom", "az_bottom",
                                  "phi_bottom_hat", "q_hat", "az_bottom_hat", "h_profile")]


class K0Step(C.Structure):
    _fields_ = [("kind", C.c_int), ("discretization", C.c_int), ("mu", C.c_double),
                ("tables", vp), ("h_bottom", vp), ("h_old", vp), ("h_new", vp), ("work", vp),
                ("diag", vp), ("residual", vp)]


class Workspace(C.Structure):
    _fields_ = [("jp", vp), ("s_bcz", vp), ("s_jz", vp), ("hat0", vp), ("hat1", vp),
                ("cscr", vp), ("partials", vp), ("metrics", vp)]


_lib = None
_load_error = None


def _declare(lib):
    P = C.POINTER
    sigs = {
        "nfgs_last_error": (C.c_char_p, []),
        "nfgs_plan_create": (C.c_int, [P(GridDesc), P(vp)]),
        "nfgs_plan_destroy": (C.c_int, [vp]),
        "nfgs_partials_count": (C.c_long, [vp]),
        "nfgs_plan_table": (vp, [vp, C.c_int]),
        "nfgs_plan_table_size": (C.c_long, [vp, C.c_int]),
        "nfgs_source_digest": (C.c_char_p, []),
        "nfgs_iteration_diagnostics_workspace_bytes": (C.c_long, [vp]),
        "nfgs_diag_source_edges": (C.c_int, [vp, vp, vp, vp]),
        "nfgs_iteration_residuals": (C.c_int, [vp] + [vp] * 8 + [vp, vp, vp]),
        "nfgs_face_fractions": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp, vp, vp, vp]),
        "nfgs_plan_set_tuning": (C.c_int, [vp, C.c_int, C.c_int]),
        "nfgs_plan_set_lateral_arithmetic": (C.c_int, [vp, C.c_int]),
        "nfgs_plan_get_lateral_arithmetic": (C.c_int, [vp]),
        "nfgs_plan_get_tuning": (C.c_int, [vp, C.c_int]),
        "nfgs_copy_plan_table": (C.c_int, [vp, C.c_int, vp, vp]),
        "nfgs_modal_robin_solve": (C.c_int, [vp, vp, vp, vp, vp, C.c_double, C.c_int, vp, vp, vp,
                                             vp]),
        "nfgs_bottom_jz": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, vp]),
        "nfgs_lap_xy": (C.c_int, [vp, vp, vp, C.c_int, vp]),
        "nfgs_reconstruct_pt": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp, vp, vp]),
        "nfgs_k0_field": (C.c_int, [vp, vp, vp, vp, vp]),
        "nfgs_transform": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_long, C.c_long, C.c_long, vp,
                                     C.c_long, C.c_long, C.c_long, C.c_int, C.c_int, C.c_int, vp]),
        "nfgs_robin_solve": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp]),
        "nfgs_dirichlet_solve": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp]),
        "nfgs_reconstruct_b": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp]),
        "nfgs_face_pass": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp, vp]),
        "nfgs_face_pass_j": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp]),
        "nfgs_state_pass": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp, vp, vp]),
        "nfgs_sources": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "nfgs_reduce": (C.c_int, [vp, vp, vp, vp]),
        "nfgs_advance": (C.c_int, [vp, P(Boundary), P(Workspace), vp, vp, vp, C.c_int, vp]),
        "nfgs_advance_ex": (C.c_int, [vp, P(Boundary), P(Workspace), vp, vp, vp, C.c_int, C.c_int,
                                      vp]),
        "nfgs_setup_jz": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, vp]),
        "nfgs_setup_poisson2d_dirichlet": (C.c_int, [vp, vp, C.c_double, vp, vp, vp]),
        "nfgs_setup_plane_lap": (C.c_int, [vp, vp, C.c_double, C.c_int, vp, vp]),
        "nfgs_gaussian_workspace_bytes": (C.c_long, [vp]),
        "nfgs_setup_gaussian": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                          vp]),
        "nfgs_project_workspace_bytes": (C.c_long, [vp]),
        "nfgs_project_field": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp, vp, vp]),
        "nfgs_above_field_workspace_bytes": (C.c_long, [vp]),
        "nfgs_gaussian_above_field": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "nfgs_sq_diff": (C.c_int, [vp, vp, vp, C.c_long, vp, vp]),
        "nfgs_fd2_curl_div": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, vp, vp, vp]),
        "nfgs_scores_workspace_bytes": (C.c_long, []),
        "nfgs_field_scores": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, vp, vp,
                                        vp]),
        "nfgs_field_scores_static_capacity": (C.c_long, []),
        "nfgs_field_scores_static": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_double, vp, vp]),
        "nfgs_outward_errors": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, P(C.c_double), C.c_int, C.c_double, vp, vp]),
        "nfgs_source_diagnostics_workspace_bytes": (C.c_long, [vp]),
        "nfgs_source_diagnostics": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp, vp, vp]),
        # coupled horizontal-harmonic closures (nfgs_k0_*)
        "nfgs_k0_table_count": (C.c_long, [C.c_int]),
        "nfgs_k0_tables": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                     vp, P(C.c_double)]),
        "nfgs_k0_update": (C.c_int, [vp, P(K0Step), vp, vp, vp, vp]),
        "nfgs_advance_k0": (C.c_int, [vp, P(Boundary), P(Workspace), vp, vp, vp, C.c_int, C.c_int,
                                      P(K0Step), vp]),
        "nfgs_k0_state_pass": (C.c_int, [vp, P(Boundary), C.c_int, vp, vp, vp, vp, vp, vp, vp,
                                         vp]),
        "nfgs_advance_relaxed": (C.c_int, [vp, P(Boundary), P(Workspace), vp, vp, C.c_double,
                                           C.c_int, vp, vp, vp, C.c_int, vp, vp]),
        # soft_cfcf start (nfgs_soft_*)
        "nfgs_soft_workspace_bytes": (C.c_long, [vp]),
        "nfgs_soft_start": (C.c_int, [vp, P(Boundary), C.c_double, C.c_double, vp, vp, vp, vp,
                                      vp]),
        "nfgs_soft_diagnostics": (C.c_int, [vp, P(Boundary), vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        # learned start (nfgs_fno_*)
        "nfgs_fno_pointwise": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, vp, vp]),
        "nfgs_fno_pointwise_backward": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int,
                                                  vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]),
        "nfgs_fno_outer_sum_workspace_bytes": (C.c_long, [C.c_int, C.c_int]),
        "nfgs_fno_outer_sum": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_long, vp, vp, C.c_int, vp,
                                         vp]),
        "nfgs_fno_axis": (C.c_int, [vp, vp, vp, C.c_long, C.c_int, C.c_long, C.c_int, vp]),
        "nfgs_fno_mix": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp]),
        "nfgs_fno_mix_grad": (C.c_int, [vp, vp, vp, C.c_int, vp]),
        "nfgs_fno_loss_workspace_bytes": (C.c_long, []),
        "nfgs_fno_loss": (C.c_int, [vp, vp, C.c_long, vp, vp, vp, vp]),
        "nfgs_fno_adam": (C.c_int, [vp, vp, vp, vp, vp, C.c_long, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_int, C.c_double, vp]),
        "nfgs_fno_to_bf16": (C.c_int, [vp, vp, C.c_long, vp]),
        "nfgs_fno_features": (C.c_int, [C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.c_double, C.c_double, vp, vp]),
        "nfgs_fno_targets": (C.c_int, [C.c_int, vp, vp, C.c_long, C.c_double, vp, vp]),
        "nfgs_fno_apply": (C.c_int, [C.c_int, vp, vp, C.c_long, C.c_double, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load():
    """Load the HIP library once; raise (never fall back) when unavailable."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m nfgs_amd.build`")
    try:
        lib = _declare(C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0)))
    except (OSError, AttributeError) as exc:  # pragma: no cover - depends on the machine
        _load_error = exc
        raise RuntimeError(f"cannot load {LIB_PATH} (rebuild it with `python -m nfgs_amd.build`): "
                           f"{exc}") from exc
    # the library must be built from the sources beside it
    from .build import source_digest
    built = lib.nfgs_source_digest().decode()
    if built != source_digest():
        raise RuntimeError(f"{LIB_PATH} was built from other sources than those in "
                           f"{LIB_PATH.parent / 'csrc'} (library {built[:12]}, sources "
                           f"{source_digest()[:12]}); rebuild it with `python -m nfgs_amd.build`")
    _lib = lib
    return _lib


class NFGSError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    """Map a C status code to the reference's exception classes."""
    if rc == NFGS_OK:
        return
    msg = load().nfgs_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == NFGS_ERR_ARG:
        raise ValueError(text)
    if rc == NFGS_ERR_STATE:
        raise RuntimeError(text)
    raise NFGSError(f"{text} (status {rc})")


def ptr(t) -> int:
    """Device address of a torch tensor (None -> NULL)."""
    return None if t is None else t.data_ptr()
