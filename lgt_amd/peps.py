"""Python host mirror of the reference's Julia calling surface over libpeps_hip.so.

Function names and argument meaning follow finite_peps_boundary_mps.jl,
finite_peps_fullupdate_v4.jl, finite_peps_gauge_invariant.jl and
finite_peps_full_update.jl; every call goes through the C ABI in
include/peps.h.  Arrays cross the boundary column-major (numpy order='F'),
complex as interleaved complex128 — the byte layout of Julia ComplexF64.

There is no CPU fallback: importing this module without a built
libpeps_hip.so raises, and every failing call raises PepsError with the
library's message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PEPS_HIP_LIB", os.path.join(_HERE, "libpeps_hip.so"))

PEPS_H, PEPS_V = 0, 1
PEPS_STEP_V4, PEPS_STEP_GAUGE, PEPS_STEP_DENSE, PEPS_STEP_FULL_V2, PEPS_STEP_REALTIME = 0, 1, 2, 3, 4
TIERS = {"fp64": 0, "fp32": 1, "bf16": 2}
STATUS = {0: "PEPS_OK", 1: "PEPS_EARG", 2: "PEPS_ESHAPE", 3: "PEPS_ENOSECTOR", 4: "PEPS_EHIP",
          5: "PEPS_ENOMEM", 6: "PEPS_ESOLVER", 7: "PEPS_EUNSUPPORTED"}

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p


class PepsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("g", C.c_double), ("t_hop", C.c_double), ("m", C.c_double),
                ("tau", C.c_double), ("D_max", C.c_int32), ("chi", C.c_int32),
                ("n_als", C.c_int32), ("flags", C.c_int32), ("als_tol", C.c_double), ("mu", C.c_double)]


# name: (restype, argtypes)
_SIGS = {
    "peps_ctx_create": (C.c_int32, [C.c_int32, C.POINTER(_vp)]),
    "peps_ctx_destroy": (C.c_int32, [_vp]),
    "peps_last_error": (C.c_char_p, [_vp]),
    "peps_warning_count": (C.c_int64, [_vp]),
    "peps_set_tier": (C.c_int32, [_vp, C.c_int32]),
    "peps_set_compress_mode": (C.c_int32, [_vp, C.c_int32]),
    "peps_set_compress_cut": (C.c_int32, [_vp, C.c_double]),
    "peps_set_solver_limits": (C.c_int32, [_vp, C.c_int32, C.c_int32]),
    "peps_set_timing": (C.c_int32, [_vp, C.c_int32]),
    "peps_stats_count": (C.c_int32, [_vp, _i32p]),
    "peps_stats_get": (C.c_int32, [_vp, C.c_int32, C.c_char_p, C.c_int32, _dp, _dp, _dp, _i64p]),
    "peps_stats_reset": (C.c_int32, [_vp]),
    "peps_sync": (C.c_int32, [_vp]),
    "peps_state_create": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _i64p, C.c_int32, C.POINTER(_vp)]),
    "peps_state_destroy": (C.c_int32, [_vp]),
    "peps_state_set_site": (C.c_int32, [_vp, C.c_int32, C.c_int32, _i64p, _dp]),
    "peps_state_site_dims": (C.c_int32, [_vp, C.c_int32, C.c_int32, _i64p]),
    "peps_state_get_site": (C.c_int32, [_vp, C.c_int32, C.c_int32, _dp]),
    "peps_state_set_bond": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int64, _dp, _i64p]),
    "peps_state_bond_dim": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _i64p]),
    "peps_state_get_bond": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, _i64p]),
    "peps_apply_mask": (C.c_int32, [_vp, C.c_int32, C.c_int32]),
    "peps_forbidden_norm": (C.c_int32, [_vp, _dp]),
    "peps_mean_nf": (C.c_int32, [_vp, _dp]),
    "peps_list_len": (C.c_int32, [_vp, _i32p]),
    "peps_list_dims": (C.c_int32, [_vp, C.c_int32, _i32p, _i64p, _i32p]),
    "peps_list_get": (C.c_int32, [_vp, C.c_int32, _dp]),
    "peps_list_destroy": (C.c_int32, [_vp]),
    "peps_dl_traced": (C.c_int32, [_vp, C.c_int32, _i64p, _dp, _dp]),
    "peps_absorb_row": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _i64p, C.POINTER(_dp), _i64p, C.POINTER(_dp), C.POINTER(_vp)]),
    "peps_compress_mps": (C.c_int32, [_vp, C.c_int32, C.c_int32, _i64p, C.POINTER(_dp), C.c_int32, C.c_double, C.POINTER(_vp)]),
    "peps_dmrg": (C.c_int32, [_vp, C.c_int32, _i64p, C.POINTER(_dp), _i64p, C.POINTER(_dp), C.c_int32, C.c_int32,
                              C.c_double, C.c_int32, C.c_int32, C.c_double, C.POINTER(C.c_double),
                              C.POINTER(C.c_int32), C.POINTER(_vp)]),
    "peps_mpo_expect": (C.c_int32, [_vp, C.c_int32, _i64p, C.POINTER(_dp), _i64p, C.POINTER(_dp),
                                    C.POINTER(C.c_double)]),
    "peps_mpo_expect_c": (C.c_int32, [_vp, C.c_int32, _i64p, C.POINTER(_dp), C.c_int32, _i64p, C.POINTER(_dp),
                                      C.POINTER(C.c_double)]),
    "peps_krylov_evolve": (C.c_int32, [_vp, C.c_int32, _i64p, C.POINTER(_dp), C.c_int32, _i64p, C.POINTER(_dp),
                                       C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.POINTER(_vp)]),
    "peps_mps_local_expect": (C.c_int32, [_vp, C.c_int32, C.c_int32, _i64p, C.POINTER(_dp), C.c_int32, _i32p,
                                          _i64p, C.POINTER(_dp), _dp]),
    "peps_boundary": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "peps_norm2": (C.c_int32, [_vp, C.c_int32, _dp]),
    "peps_bond_env": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "peps_charge_svd": (C.c_int32, [_vp, C.c_int32, C.c_int64, C.c_int64, _dp, _i64p, _i64p, C.c_int32, C.c_double, _i64p, _i64p, C.POINTER(_vp)]),
    "peps_als_truncate": (C.c_int32, [_vp, C.c_int32, _i64p, _dp, _dp, C.c_int32, _i64p, C.c_int32, C.c_double, _i64p, _i64p, _dp, C.POINTER(_vp)]),
    "peps_als_truncate_init": (C.c_int32, [_vp, C.c_int32, _i64p, _dp, _dp, C.c_int32, _i64p, C.c_int32, C.c_double,
                                           C.c_double, _dp, _dp, _i64p, _i64p, _dp, C.POINTER(_vp)]),
    "peps_su_bond": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, C.c_int32]),
    "peps_fu_bond_v4": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, C.c_int32, C.c_int32, C.c_int32]),
    "peps_trotter_step": (C.c_int32, [_vp, C.c_int32, C.POINTER(Params)]),
    "peps_batch_trotter_step": (C.c_int32, [C.POINTER(_vp), C.c_int32, C.c_int32, C.POINTER(Params), _i32p]),
    "peps_last_als_stats": (C.c_int32, [_vp, _i64p]),
    "peps_apply_site_op": (C.c_int32, [_vp, C.c_int32, C.c_int32, _dp]),
    "peps_ite_ground_state": (C.c_int32, [_vp, C.c_int32, C.POINTER(Params), C.c_int64, _i64p]),
    "peps_dl_open": (C.c_int32, [_vp, C.c_int32, _i64p, _dp, _dp]),
    "peps_energy_v4": (C.c_int32, [_vp, C.POINTER(Params), _dp]),
    "peps_su_bond_dense": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, C.c_int32]),
    "peps_fu_bond_v2": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, C.c_int32, C.c_int32]),
    "peps_plaquette_gate": (C.c_int32, [_vp, C.c_int32, C.c_double, C.c_double, C.c_double, _i32p, _i32p,
                                        _dp, _dp, _dp]),
    "peps_update_plaquette": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp,
                                          C.c_int32]),
    "peps_expect_site": (C.c_int32, [_vp, C.c_int32, C.c_int32, _dp, _dp]),
    "peps_plaquette_energy": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_double, _dp]),
    "peps_total_energy": (C.c_int32, [_vp, C.POINTER(Params), _dp]),
    "peps_plaquette_expect": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp]),
    "peps_bond_entropy": (C.c_int32, [_dp, C.c_int64, _dp]),
    "peps_measure_all": (C.c_int32, [_vp, _dp, _dp, _dp]),
    "peps_mlp_create": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(_vp)]),
    "peps_mlp_destroy": (C.c_int32, [_vp]),
    "peps_mlp_param_count": (C.c_int32, [_vp, _i64p]),
    "peps_mlp_get_params": (C.c_int32, [_vp, _dp]),
    "peps_mlp_set_params": (C.c_int32, [_vp, _dp]),
    "peps_mlp_forward": (C.c_int32, [_vp, C.c_int64, _dp, _dp]),
    "peps_mlp_gradients": (C.c_int32, [_vp, C.c_int64, _dp, _dp, _dp, _dp]),
    "peps_mlp_train_step": (C.c_int32, [_vp, C.c_int64, _dp, _dp, C.c_double, _dp]),
    "peps_gemm": (C.c_int32, [_vp, C.c_int32, C.c_int64, C.c_int64, C.c_int64, _dp, C.c_int32, C.c_int32, _dp, C.c_int32, C.c_int32, _dp]),
    "peps_ed_create": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _i64p, C.POINTER(_vp)]),
    "peps_ed_destroy": (C.c_int32, [_vp]),
    "peps_ed_dim": (C.c_int32, [_vp, _i64p]),
    "peps_ed_keys": (C.c_int32, [_vp, C.POINTER(C.c_uint64)]),
    "peps_ed_hamiltonian": (C.c_int32, [_vp, C.c_double, C.c_double, C.c_double]),
    "peps_ed_apply": (C.c_int32, [_vp, _dp, _dp]),
    "peps_ed_ground_state": (C.c_int32, [_vp, C.c_double, _dp, _dp]),
    "peps_ed_evolve": (C.c_int32, [_vp, _dp, C.c_double, C.c_int32, C.c_int32, C.c_double, _dp, _dp, _dp, _dp,
                                   _i32p]),
    "peps_ed_measure": (C.c_int32, [_vp, _dp, _dp, _dp, _dp]),
    "peps_gemm_noalloc": (C.c_int32, [_vp, C.c_int32, C.c_int64, C.c_int64, C.c_int64, _vp, C.c_int32, C.c_int32,
                                      _vp, C.c_int32, C.c_int32, _vp]),
    "peps_permute": (C.c_int32, [_vp, C.c_int32, C.c_int32, _i64p, _i32p, _dp, _dp]),
    "peps_gemm_acc": (C.c_int32, [_vp, C.c_int32, C.c_int64, C.c_int64, C.c_int64, _dp, C.c_int32, C.c_int32, _dp,
                                  C.c_int32, C.c_int32, C.c_double, _dp]),
    "peps_eigh": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp]),
    "peps_solve_herm": (C.c_int32, [_vp, C.c_int32, C.c_int64, C.c_int64, _dp, _dp, C.c_double, _dp]),
    "peps_svd": (C.c_int32, [_vp, C.c_int32, C.c_int64, C.c_int64, _dp, _dp, _dp, _dp]),
    "peps_qr": (C.c_int32, [_vp, C.c_int32, C.c_int64, C.c_int64, _dp, _dp, _dp]),
    "peps_bond_operator": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_double, C.c_double, C.c_double, C.c_double, _i32p, _dp]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """Load libpeps_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libpeps_hip.so not found at {LIB_PATH}: run `python -m lgt_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_i64p)


class Context:
    """peps_ctx: one HIP stream, a caching allocator and the precision tier."""

    def __init__(self, device: int = 0):
        h = _vp()
        st = lib().peps_ctx_create(device, C.byref(h))
        if st != 0:
            raise PepsError(st, "peps_ctx_create failed")
        self.h = h

    def check(self, st):
        if st != 0:
            raise PepsError(st, lib().peps_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().peps_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tier(self, tier: str):
        self.check(lib().peps_set_tier(self.h, TIERS[tier]))

    def set_compress_mode(self, mode: str):
        """'auto', 'qr_svd' (explicit product) or 'gram' (factored Gram form)."""
        self.check(lib().peps_set_compress_mode(self.h, {"auto": 0, "qr_svd": 1, "gram": 2}[mode]))

    def set_compress_cut(self, cut: float):
        """Singular-value cut of the boundary compression (1e-14 by default)."""
        self.check(lib().peps_set_compress_cut(self.h, float(cut)))

    def set_timing(self, on, gemm_only: bool = False):
        """Per-kernel-class HIP-event timing; gemm_only times just the MFMA
        GEMM classes of the tiers (the event pairs of every class cost about
        4% of a full-update step)."""
        self.check(lib().peps_set_timing(self.h, 0 if not on else (2 if gemm_only else 1)))

    def sync(self):
        self.check(lib().peps_sync(self.h))

    def set_solver_limits(self, subspace_max_it: int = 25, jacobi_max_sweeps: int = 80):
        """Iteration caps of the subspace iteration and the Jacobi SVD; a cap
        reached counts a warning."""
        self.check(lib().peps_set_solver_limits(self.h, int(subspace_max_it), int(jacobi_max_sweeps)))

    def warning_count(self) -> int:
        return int(lib().peps_warning_count(self.h))

    def stats(self) -> dict:
        n = C.c_int32()
        self.check(lib().peps_stats_count(self.h, C.byref(n)))
        out = {}
        for i in range(n.value):
            name = C.create_string_buffer(64)
            f, b, ms = C.c_double(), C.c_double(), C.c_double()
            k = C.c_int64()
            self.check(lib().peps_stats_get(self.h, i, name, 64, C.byref(f), C.byref(b), C.byref(ms), C.byref(k)))
            out[name.value.decode()] = dict(flops=f.value, bytes=b.value, ms=ms.value, launches=k.value)
        return out

    def stats_reset(self):
        self.check(lib().peps_stats_reset(self.h))


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _as_array(a, cplx):
    if cplx:
        return np.asfortranarray(a, dtype=np.complex128)
    a = np.asarray(a)
    if np.iscomplexobj(a):
        if np.any(a.imag != 0):
            raise PepsError(1, "complex data passed to a real state/context call")
        a = a.real
    return np.asfortranarray(a, dtype=np.float64)


def _take_list(ctx: Context, h) -> list:
    L = lib()
    n = C.c_int32()
    ctx.check(L.peps_list_len(h, C.byref(n)))
    out = []
    try:
        for i in range(n.value):
            rank = C.c_int32()
            dims = (C.c_int64 * 8)()
            cp = C.c_int32()
            ctx.check(L.peps_list_dims(h, i, C.byref(rank), dims, C.byref(cp)))
            shp = tuple(dims[k] for k in range(rank.value))
            a = np.empty(shp, dtype=np.complex128 if cp.value else np.float64, order="F")
            ctx.check(L.peps_list_get(h, i, a.ctypes.data_as(_dp)))
            out.append(a)
    finally:
        L.peps_list_destroy(h)
    return out


# ════════════════════════════════════════════════════════════════════════════
#  GaugePEPS on the device
# ════════════════════════════════════════════════════════════════════════════

class GaugePEPS:
    """Device-resident GaugePEPS (finite_peps_gauge_invariant.jl:82-88).

    ``tensors[ix-1][iy-1]`` (p,l,r,u,d); ``lh[ix-1][iy-1]``, ``lv[...]`` weights;
    ``qh``, ``qv`` integer flux labels (2E); ``gch`` (nx, ny) background charges.
    """

    def __init__(self, nx, ny, gd, gch, is_complex=False, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.nx, self.ny, self.gd = nx, ny, gd
        self.is_complex = bool(is_complex)
        g = np.asfortranarray(np.asarray(gch, dtype=np.int64).reshape(nx, ny))
        self.gch = g
        h = _vp()
        self.ctx.check(lib().peps_state_create(self.ctx.h, nx, ny, gd, _iptr(g), int(self.is_complex), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().peps_state_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # ── upload / download ────────────────────────────────────────────────────
    @classmethod
    def from_arrays(cls, tensors, lh, lv, qh, qv, gch, gd, is_complex=False, ctx=None):
        nx, ny = len(tensors), len(tensors[0])
        gp = cls(nx, ny, gd, gch, is_complex, ctx)
        for ix in range(1, nx + 1):
            for iy in range(1, ny + 1):
                gp.set_site(ix, iy, tensors[ix - 1][iy - 1])
        for iy in range(1, ny + 1):
            for ix in range(1, nx):
                gp.set_bond(PEPS_H, ix, iy, lh[ix - 1][iy - 1], qh[ix - 1][iy - 1])
        for iy in range(1, ny):
            for ix in range(1, nx + 1):
                gp.set_bond(PEPS_V, ix, iy, lv[ix - 1][iy - 1], qv[ix - 1][iy - 1])
        return gp

    def set_site(self, ix, iy, T):
        a = _as_array(T, self.is_complex)
        assert a.ndim == 5
        d = np.array(a.shape, dtype=np.int64)
        self.ctx.check(lib().peps_state_set_site(self.h, ix, iy, _iptr(d), _dptr(a)))

    def site(self, ix, iy):
        d = np.zeros(5, np.int64)
        self.ctx.check(lib().peps_state_site_dims(self.h, ix, iy, _iptr(d)))
        a = np.empty(tuple(d), dtype=np.complex128 if self.is_complex else np.float64, order="F")
        self.ctx.check(lib().peps_state_get_site(self.h, ix, iy, _dptr(a)))
        return a

    def set_bond(self, orient, ix, iy, lam, flux):
        l = np.ascontiguousarray(lam, dtype=np.float64)
        q = np.ascontiguousarray(flux, dtype=np.int64)
        assert len(l) == len(q)
        self.ctx.check(lib().peps_state_set_bond(self.h, orient, ix, iy, len(l), _dptr(l), _iptr(q)))

    def bond(self, orient, ix, iy):
        D = C.c_int64()
        self.ctx.check(lib().peps_state_bond_dim(self.h, orient, ix, iy, C.byref(D)))
        l = np.zeros(D.value)
        q = np.zeros(D.value, np.int64)
        self.ctx.check(lib().peps_state_get_bond(self.h, orient, ix, iy, _dptr(l), _iptr(q)))
        return l, q

    def to_arrays(self):
        tens = [[self.site(ix, iy) for iy in range(1, self.ny + 1)] for ix in range(1, self.nx + 1)]
        lh = [[None] * self.ny for _ in range(self.nx - 1)]
        qh = [[None] * self.ny for _ in range(self.nx - 1)]
        lv = [[None] * (self.ny - 1) for _ in range(self.nx)]
        qv = [[None] * (self.ny - 1) for _ in range(self.nx)]
        for iy in range(1, self.ny + 1):
            for ix in range(1, self.nx):
                lh[ix - 1][iy - 1], qh[ix - 1][iy - 1] = self.bond(PEPS_H, ix, iy)
        for iy in range(1, self.ny):
            for ix in range(1, self.nx + 1):
                lv[ix - 1][iy - 1], qv[ix - 1][iy - 1] = self.bond(PEPS_V, ix, iy)
        return dict(tensors=tens, lh=lh, lv=lv, qh=qh, qv=qv)


PEPS_STEP_RENORM = 1
PEPS_STEP_POSITIVE_COST = 2


def _params(g, t_hop, m, tau, D_max, chi, n_als, als_tol=0.0, flags=0, mu=0.0):
    return Params(g=g, t_hop=t_hop, m=m, tau=tau, D_max=D_max, chi=chi, n_als=n_als, flags=flags,
                  als_tol=als_tol, mu=mu)


# ── finite_peps_gauge_invariant.jl ───────────────────────────────────────────

def apply_mask(gp: GaugePEPS, ix, iy):
    gp.ctx.check(lib().peps_apply_mask(gp.h, ix, iy))


def forbidden_norm(gp: GaugePEPS) -> float:
    v = C.c_double()
    gp.ctx.check(lib().peps_forbidden_norm(gp.h, C.byref(v)))
    return v.value


def mean_nf(gp: GaugePEPS) -> float:
    v = C.c_double()
    gp.ctx.check(lib().peps_mean_nf(gp.h, C.byref(v)))
    return v.value


def charge_svd(M, row_charge, col_charge, D_max, reg=1e-12, ctx=None):
    """charge_svd (finite_peps_gauge_invariant.jl:236-281) -> (U, S, Vt, bond_charge)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(M)
    a = _as_array(M, cp)
    rq = np.ascontiguousarray(row_charge, dtype=np.int64)
    cq = np.ascontiguousarray(col_charge, dtype=np.int64)
    Dk = C.c_int64()
    q = np.zeros(D_max, np.int64)
    h = _vp()
    ctx.check(lib().peps_charge_svd(ctx.h, int(cp), a.shape[0], a.shape[1], _dptr(a), _iptr(rq), _iptr(cq),
                                    D_max, reg, C.byref(Dk), _iptr(q), C.byref(h)))
    U, Vt, S = _take_list(ctx, h)
    return U, S, Vt, q[:Dk.value]


def simple_update_bond_h(gp: GaugePEPS, ix, iy, gate, D_max):
    """simple_update_bond_h! (finite_peps_gauge_invariant.jl:296)."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_su_bond(gp.h, PEPS_H, ix, iy, _dptr(G), D_max))


def simple_update_bond_v(gp: GaugePEPS, ix, iy, gate, D_max):
    """simple_update_bond_v! (finite_peps_gauge_invariant.jl:412)."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_su_bond(gp.h, PEPS_V, ix, iy, _dptr(G), D_max))


def trotter_step_gauge(gp: GaugePEPS, g, t_hop, m, tau, D_max):
    """trotter_step_gauge! (finite_peps_gauge_invariant.jl:523)."""
    p = _params(g, t_hop, m, tau, D_max, 1, 0)
    gp.ctx.check(lib().peps_trotter_step(gp.h, PEPS_STEP_GAUGE, C.byref(p)))


# ── finite_peps_ground_state.jl (dense simple update with plaquettes) ───────
# These treat the state as a FinitePEPS (no mask, no flux labels): after a
# dense update the gauge-invariant calls refuse the state.

def update_bond_h(gp: GaugePEPS, ix, iy, gate, D_trunc):
    """update_bond_h! (finite_peps_ground_state.jl:175)."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_su_bond_dense(gp.h, PEPS_H, ix, iy, _dptr(G), D_trunc))


def update_bond_v(gp: GaugePEPS, ix, iy, gate, D_trunc):
    """update_bond_v! (finite_peps_ground_state.jl:280)."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_su_bond_dense(gp.h, PEPS_V, ix, iy, _dptr(G), D_trunc))


def full_update_bond_h(gp: GaugePEPS, ix, iy, gate, D_trunc, use_gauss=False):
    """full_update_bond_h! (finite_peps_full_update.jl:346); use_gauss=True is refused."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_fu_bond_v2(gp.h, PEPS_H, ix, iy, _dptr(G), D_trunc, int(use_gauss)))


def full_update_bond_v(gp: GaugePEPS, ix, iy, gate, D_trunc, use_gauss=False):
    """full_update_bond_v! (finite_peps_full_update.jl:462)."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_fu_bond_v2(gp.h, PEPS_V, ix, iy, _dptr(G), D_trunc, int(use_gauss)))


def trotter_step_full(gp: GaugePEPS, g, t_hop, m, tau, D_trunc):
    """trotter_step_full! (finite_peps_full_update.jl:550), use_gauss=false."""
    p = _params(g, t_hop, m, tau, D_trunc, 1, 0)
    gp.ctx.check(lib().peps_trotter_step(gp.h, PEPS_STEP_FULL_V2, C.byref(p)))


def plaquette_gate(gd, g, tau, cutoff=1e-12, ctx=None):
    """decompose_plaquette_LShape(exp(-tau H_plaquette_gauge)) (u1_lgt_hamiltonian.jl:728)
    -> (GA, GB, GC)."""
    ctx = ctx or default_context()
    Da, Db = C.c_int32(), C.c_int32()
    nul = C.cast(None, _dp)
    ctx.check(lib().peps_plaquette_gate(ctx.h, gd, g, tau, cutoff, C.byref(Da), C.byref(Db), nul, nul, nul))
    GA = np.zeros((gd, gd, gd, gd, Da.value, Db.value), order="F")
    GB = np.zeros((gd, gd, Da.value), order="F")
    GC = np.zeros((gd, gd, Db.value), order="F")
    ctx.check(lib().peps_plaquette_gate(ctx.h, gd, g, tau, cutoff, C.byref(Da), C.byref(Db), _dptr(GA),
                                        _dptr(GB), _dptr(GC)))
    return GA, GB, GC


def update_plaquette(gp: GaugePEPS, ix, iy, GA, GB, GC, D_trunc):
    """update_plaquette! (finite_peps_ground_state.jl:405)."""
    A, B, Cc = (_as_array(x, gp.is_complex) for x in (GA, GB, GC))
    gp.ctx.check(lib().peps_update_plaquette(gp.h, ix, iy, A.shape[4], A.shape[5], _dptr(A), _dptr(B), _dptr(Cc),
                                             D_trunc))


def trotter_step(gp: GaugePEPS, g, t_hop, m, tau, D_trunc, mu=0.0):
    """trotter_step! (finite_peps_ground_state.jl:498): h-half, v-half, plaquettes, v-half, h-half."""
    p = _params(g, t_hop, m, tau, D_trunc, 1, 0, mu=mu)
    gp.ctx.check(lib().peps_trotter_step(gp.h, PEPS_STEP_DENSE, C.byref(p)))


def realtime_trotter_step(gp: GaugePEPS, g, t_hop, m, dt, D_trunc):
    """realtime_trotter_step! (finite_peps_quench.jl:81-109): exp(-i dt/2 H_h),
    exp(-i dt H_v), exp(-i dt/2 H_h) through the dense bond updates; the
    state must be complex."""
    p = _params(g, t_hop, m, dt, D_trunc, 1, 0)
    gp.ctx.check(lib().peps_trotter_step(gp.h, PEPS_STEP_REALTIME, C.byref(p)))


def apply_site_op(gp: GaugePEPS, ix, iy, O):
    """apply_site_op! (finite_peps_quench.jl:228-234): T(ix,iy) <- O T on the
    physical index; the flux labels are undefined afterwards."""
    a = _as_array(O, gp.is_complex)
    gp.ctx.check(lib().peps_apply_site_op(gp.h, ix, iy, _dptr(a)))


def expect_site(gp: GaugePEPS, ix, iy, O):
    """expect_site (finite_peps_ground_state.jl:549)."""
    a = _as_array(O, gp.is_complex)
    v = C.c_double()
    gp.ctx.check(lib().peps_expect_site(gp.h, ix, iy, _dptr(a), C.byref(v)))
    return v.value


def plaquette_energy(gp: GaugePEPS, ix, iy, g):
    """plaquette_energy (finite_peps_ground_state.jl:589)."""
    v = C.c_double()
    gp.ctx.check(lib().peps_plaquette_energy(gp.h, ix, iy, g, C.byref(v)))
    return v.value


def total_energy(gp: GaugePEPS, g, t_hop, m):
    """total_energy (finite_peps_ground_state.jl:690)."""
    p = _params(g, t_hop, m, 0.0, 1, 1, 0)
    v = C.c_double()
    gp.ctx.check(lib().peps_total_energy(gp.h, C.byref(p), C.byref(v)))
    return v.value


# ── finite_peps_quench.jl (measurements) ────────────────────────────────────

def plaquette_expect(gp: GaugePEPS, ix, iy, chi) -> complex:
    """<W> on the plaquette with lower-left site (ix, iy) from boundary MPS."""
    out = np.zeros(2)
    gp.ctx.check(lib().peps_plaquette_expect(gp.h, ix, iy, chi, _dptr(out)))
    return complex(out[0], out[1])


def bond_entropy(lam) -> float:
    """bond_entropy (finite_peps_quench.jl:121)."""
    a = np.ascontiguousarray(lam, dtype=np.float64)
    v = C.c_double()
    rc = lib().peps_bond_entropy(_dptr(a), len(a), C.byref(v))
    if rc:
        raise PepsError(rc, "bond_entropy")
    return v.value


def measure_all_finite(gp: GaugePEPS, t=0.0) -> dict:
    """measure_all_finite (finite_peps_quench.jl:166)."""
    out = np.zeros(7)
    nf = np.zeros((gp.nx, gp.ny), order="F")
    e2 = np.zeros((gp.nx, gp.ny), order="F")
    gp.ctx.check(lib().peps_measure_all(gp.h, _dptr(out), _dptr(nf), _dptr(e2)))
    keys = ("nf_mean", "nf_even", "nf_odd", "E2_mean", "S_h_mean", "S_v_mean", "S_mean")
    return dict(t=t, **dict(zip(keys, out.tolist())), nf_grid=nf, E2_grid=e2)


# ── learned ALS warm start: the GPU MLP ─────────────────────────────────────

class Mlp:
    """in -> width -> width -> out perceptron (ReLU hidden layers) trained and
    evaluated on the GPU; batch along columns (numpy arrays (n, batch))."""

    def __init__(self, n_in, width, n_out, seed=0, ctx=None):
        self.ctx = ctx or default_context()
        self.n_in, self.width, self.n_out = n_in, width, n_out
        h = _vp()
        self.ctx.check(lib().peps_mlp_create(self.ctx.h, n_in, width, n_out, seed, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().peps_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def param_count(self):
        n = C.c_int64()
        self.ctx.check(lib().peps_mlp_param_count(self.h, C.byref(n)))
        return n.value

    def get_params(self):
        p = np.zeros(self.param_count())
        self.ctx.check(lib().peps_mlp_get_params(self.h, _dptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.size == self.param_count()
        self.ctx.check(lib().peps_mlp_set_params(self.h, _dptr(p)))

    def forward(self, X):
        X = np.asfortranarray(X, dtype=np.float64)
        Y = np.zeros((self.n_out, X.shape[1]), order="F")
        self.ctx.check(lib().peps_mlp_forward(self.h, X.shape[1], _dptr(X), _dptr(Y)))
        return Y

    def gradients(self, X, T):
        X = np.asfortranarray(X, dtype=np.float64)
        T = np.asfortranarray(T, dtype=np.float64)
        g = np.zeros(self.param_count())
        loss = C.c_double()
        self.ctx.check(lib().peps_mlp_gradients(self.h, X.shape[1], _dptr(X), _dptr(T), C.byref(loss), _dptr(g)))
        return loss.value, g

    def train_step(self, X, T, lr=1e-3):
        X = np.asfortranarray(X, dtype=np.float64)
        T = np.asfortranarray(T, dtype=np.float64)
        loss = C.c_double()
        self.ctx.check(lib().peps_mlp_train_step(self.h, X.shape[1], _dptr(X), _dptr(T), lr, C.byref(loss)))
        return loss.value


# ── finite_peps_full_update.jl / finite_peps_boundary_mps.jl ────────────────

def dl_traced(T, ctx=None):
    """_dl_traced (finite_peps_full_update.jl:79-91)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(T)
    a = _as_array(T, cp)
    d = np.array(a.shape, dtype=np.int64)
    E = np.empty((d[1] ** 2, d[2] ** 2, d[3] ** 2, d[4] ** 2), dtype=a.dtype, order="F")
    ctx.check(lib().peps_dl_traced(ctx.h, int(cp), _iptr(d), _dptr(a), _dptr(E)))
    return E


def dl_open(T, ctx=None):
    """_dl_open (finite_peps_full_update.jl:97-109): (p, p', l², r², u², d²)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(T)
    a = _as_array(T, cp)
    d = np.array(a.shape, dtype=np.int64)
    E = np.empty((d[0], d[0], d[1] ** 2, d[2] ** 2, d[3] ** 2, d[4] ** 2), dtype=a.dtype, order="F")
    ctx.check(lib().peps_dl_open(ctx.h, int(cp), _iptr(d), _dptr(a), _dptr(E)))
    return E


def _pack(arrs, cp):
    arrs = [_as_array(x, cp) for x in arrs]
    dims = np.array([s for x in arrs for s in x.shape], dtype=np.int64)
    ptrs = (_dp * len(arrs))(*[_dptr(x) for x in arrs])
    return arrs, dims, ptrs


def _absorb(mps, row, from_top, ctx=None):
    ctx = ctx or default_context()
    cp = any(np.iscomplexobj(x) for x in list(mps) + list(row))
    ma, md, mp = _pack(mps, cp)
    ra, rd, rp = _pack(row, cp)
    h = _vp()
    ctx.check(lib().peps_absorb_row(ctx.h, int(from_top), int(cp), len(ma), _iptr(md), mp, _iptr(rd), rp, C.byref(h)))
    return _take_list(ctx, h)


def absorb_row_from_top(mps, row, ctx=None):
    """absorb_row_from_top (finite_peps_boundary_mps.jl:128)."""
    return _absorb(mps, row, True, ctx)


def absorb_row_from_bottom(mps, row, ctx=None):
    """absorb_row_from_bottom (finite_peps_boundary_mps.jl:156)."""
    return _absorb(mps, row, False, ctx)


def compress_mps(mps, chi, cutoff=1e-14, ctx=None):
    """compress_mps! (finite_peps_boundary_mps.jl:70) — returns the new list."""
    ctx = ctx or default_context()
    cp = any(np.iscomplexobj(x) for x in mps)
    a, d, p = _pack(mps, cp)
    h = _vp()
    ctx.check(lib().peps_compress_mps(ctx.h, int(cp), len(a), _iptr(d), p, chi, cutoff, C.byref(h)))
    return _take_list(ctx, h)


def dmrg(mpo, psi0, D, nsweeps=8, tol=1e-7, maxiter=60, krylovdim=12, etol=1e-7, ctx=None):
    """dmrg_ground_state (mps_lgt.jl:418-467) on the GPU for a real MPO
    (tensors (Dl, Dr, d, d)) from the start MPS psi0 (tensors (Dl, Dr, d)).
    Returns (energy, MPS, sweeps run)."""
    ctx = ctx or default_context()
    wa, wd, wp = _pack(mpo, False)
    pa, pd, pp = _pack(psi0, False)
    E, sw, h = C.c_double(), C.c_int32(), _vp()
    ctx.check(lib().peps_dmrg(ctx.h, len(wa), _iptr(wd), wp, _iptr(pd), pp, int(D), int(nsweeps), float(tol),
                              int(maxiter), int(krylovdim), float(etol), C.byref(E), C.byref(sw), C.byref(h)))
    return E.value, _take_list(ctx, h), sw.value


def _is_complex(arrs):
    return any(np.iscomplexobj(a) for a in arrs)


def mpo_expect(mpo, psi, ctx=None):
    """mpo_expect (mps_lgt.jl:308-315): <psi|W|psi> on the GPU for a real
    MPO; the real part for a complex MPS."""
    ctx = ctx or default_context()
    wa, wd, wp = _pack(mpo, False)
    cp = _is_complex(psi)
    pa, pd, pp = _pack(psi, cp)
    out = C.c_double()
    if cp:
        ctx.check(lib().peps_mpo_expect_c(ctx.h, len(wa), _iptr(wd), wp, 1, _iptr(pd), pp, C.byref(out)))
    else:
        ctx.check(lib().peps_mpo_expect(ctx.h, len(wa), _iptr(wd), wp, _iptr(pd), pp, C.byref(out)))
    return out.value


def krylov_evolve(mpo, psi0, dt, nsteps=1, m=8, Dmax=100, eps=1e-10, ctx=None):
    """krylov_evolve_step (mps_lgt.jl:546-586) nsteps times on the GPU:
    psi <- exp(-i dt H) psi with m Krylov vectors, bonds <= Dmax, relative
    cut eps, renormalised each step.  Returns the complex MPS."""
    ctx = ctx or default_context()
    wa, wd, wp = _pack(mpo, False)
    cp = _is_complex(psi0)
    pa, pd, pp = _pack(psi0, cp)
    h = _vp()
    ctx.check(lib().peps_krylov_evolve(ctx.h, len(wa), _iptr(wd), wp, int(cp), _iptr(pd), pp, float(dt), int(nsteps),
                                       int(m), int(Dmax), float(eps), C.byref(h)))
    return _take_list(ctx, h)


def mps_local_expect_many(psi, sites, ops, ctx=None):
    """mps_local_expect (mps_lgt.jl:317-330) for every (site, operator) pair
    (sites 0-based chain positions, real (d, d) operators) in one pass of
    overlap environments: Re <psi|O|psi> / <psi|psi>."""
    ctx = ctx or default_context()
    cp = _is_complex(psi)
    pa, pd, pp = _pack(psi, cp)
    oa, od, op = _pack(ops, False)
    st = np.ascontiguousarray(sites, dtype=np.int32)
    out = np.zeros(len(oa))
    if len(oa) != len(st):
        raise ValueError("mps_local_expect_many: one site per operator")
    ctx.check(lib().peps_mps_local_expect(ctx.h, len(pa), int(cp), _iptr(pd), pp, len(oa),
                                          st.ctypes.data_as(_i32p), _iptr(od), op, _dptr(out)))
    return out


def boundary_top(gp: GaugePEPS, stop_row, chi):
    """boundary_top (finite_peps_boundary_mps.jl:183)."""
    h = _vp()
    gp.ctx.check(lib().peps_boundary(gp.h, 1, stop_row, chi, C.byref(h)))
    return _take_list(gp.ctx, h)


def boundary_bottom(gp: GaugePEPS, stop_row, chi):
    """boundary_bottom (finite_peps_boundary_mps.jl:207)."""
    h = _vp()
    gp.ctx.check(lib().peps_boundary(gp.h, 0, stop_row, chi, C.byref(h)))
    return _take_list(gp.ctx, h)


def peps_norm2(gp: GaugePEPS, chi=64) -> complex:
    """peps_norm2 (finite_peps_boundary_mps.jl:228)."""
    out = np.zeros(2)
    gp.ctx.check(lib().peps_norm2(gp.h, chi, _dptr(out)))
    return complex(out[0], out[1])


def _env(gp, orient, ix, iy, chi):
    h = _vp()
    gp.ctx.check(lib().peps_bond_env(gp.h, orient, ix, iy, chi, C.byref(h)))
    env, RA, RB, QA, QB = _take_list(gp.ctx, h)
    return dict(env=env, RA=RA, RB=RB, QA=QA, QB=QB)


def bond_environment_h(gp: GaugePEPS, ix, iy, chi=64):
    """bond_environment_h (finite_peps_boundary_mps.jl:389)."""
    return _env(gp, PEPS_H, ix, iy, chi)


def bond_environment_v(gp: GaugePEPS, ix, iy, chi=64):
    """bond_environment_v (finite_peps_boundary_mps.jl:507)."""
    return _env(gp, PEPS_V, ix, iy, chi)


# ── finite_peps_fullupdate_v4.jl ─────────────────────────────────────────────

def als_truncate(theta, env, D_max, eRofpL, n_als=30, reg=1e-10, ctx=None, return_info=False):
    """als_truncate (finite_peps_fullupdate_v4.jl:94) -> (X, Y, bflux)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(theta) or np.iscomplexobj(env)
    th = _as_array(theta, cp)
    en = _as_array(env, cp)
    d = np.array(th.shape, dtype=np.int64)
    eR = np.ascontiguousarray(eRofpL, dtype=np.int64)
    Dp = C.c_int64()
    bf = np.zeros(D_max, np.int64)
    info = np.zeros(3)
    h = _vp()
    ctx.check(lib().peps_als_truncate(ctx.h, int(cp), _iptr(d), _dptr(th), _dptr(en), D_max, _iptr(eR), n_als,
                                      reg, C.byref(Dp), _iptr(bf), _dptr(info), C.byref(h)))
    X, Y = _take_list(ctx, h)
    out = (X, Y, bf[:Dp.value])
    if return_info:
        return out + (dict(cost0=info[0], cost1=info[1], accepted=bool(info[2])),)
    return out


def als_truncate_init(theta, env, D_max, eRofpL, n_als=30, reg=1e-10, als_tol=0.0, init=None, ctx=None):
    """als_truncate with an optional initial bond (X0, Y0) shaped like the SVD
    initialisation and an optional tolerance -> (X, Y, bflux, info)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(theta) or np.iscomplexobj(env) or (init is not None and any(np.iscomplexobj(a) for a in init))
    th = _as_array(theta, cp)
    en = _as_array(env, cp)
    d = np.array(th.shape, dtype=np.int64)
    eR = np.ascontiguousarray(eRofpL, dtype=np.int64)
    Dp = C.c_int64()
    bf = np.zeros(D_max, np.int64)
    info = np.zeros(4)
    h = _vp()
    if init is None:
        x0 = y0 = C.cast(None, _dp)
    else:
        X0, Y0 = (_as_array(a, cp) for a in init)
        x0, y0 = _dptr(X0), _dptr(Y0)
    ctx.check(lib().peps_als_truncate_init(ctx.h, int(cp), _iptr(d), _dptr(th), _dptr(en), D_max, _iptr(eR), n_als,
                                           reg, als_tol, x0, y0, C.byref(Dp), _iptr(bf), _dptr(info), C.byref(h)))
    X, Y = _take_list(ctx, h)
    return X, Y, bf[:Dp.value], dict(cost0=info[0], cost1=info[1], accepted=bool(info[2]), sweeps=int(info[3]))


def full_update_bond_h_v4(gp: GaugePEPS, ix, iy, gate, D_max, chi=64, n_als=30):
    """full_update_bond_h_v4! (finite_peps_fullupdate_v4.jl:165)."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_fu_bond_v4(gp.h, PEPS_H, ix, iy, _dptr(G), D_max, chi, n_als))


def full_update_bond_v_v4(gp: GaugePEPS, ix, iy, gate, D_max, chi=64, n_als=30):
    """full_update_bond_v_v4! (finite_peps_fullupdate_v4.jl:274)."""
    G = _as_array(gate, gp.is_complex)
    gp.ctx.check(lib().peps_fu_bond_v4(gp.h, PEPS_V, ix, iy, _dptr(G), D_max, chi, n_als))


def trotter_step_v4(gp: GaugePEPS, g, t_hop, m, tau, D_max, chi=64, n_als=30, als_tol=0.0, renorm=False,
                    positive_cost=False):
    """trotter_step_v4! (finite_peps_fullupdate_v4.jl:415).  Returns ALS stats.
    Options without a reference counterpart (both off by default):
    renorm normalises every boundary MPS after each row absorption
    (PEPS_STEP_RENORM: keeps the absolute truncation thresholds meaningful
    when <psi|psi> drifts far from 1); positive_cost accepts an ALS
    refinement only if its cost is >= 0 (PEPS_STEP_POSITIVE_COST: guards
    against environments that are not positive semidefinite)."""
    flags = (PEPS_STEP_RENORM if renorm else 0) | (PEPS_STEP_POSITIVE_COST if positive_cost else 0)
    p = _params(g, t_hop, m, tau, D_max, chi, n_als, als_tol, flags)
    gp.ctx.check(lib().peps_trotter_step(gp.h, PEPS_STEP_V4, C.byref(p)))
    st = np.zeros(3, np.int64)
    gp.ctx.check(lib().peps_last_als_stats(gp.h, _iptr(st)))
    return dict(bonds=int(st[0]), accepted=int(st[1]), sweeps=int(st[2]))


def trotter_step_batch(states, kind, g, t_hop, m, tau, D_max, chi=64, n_als=30, als_tol=0.0, renorm=False,
                       positive_cost=False):
    """One Trotter step (kind PEPS_STEP_V4 / _GAUGE / _DENSE / _FULL_V2) on
    each independent state at once, every state on its own Context.  Returns
    the per-state status codes; raises on the first failure."""
    flags = (PEPS_STEP_RENORM if renorm else 0) | (PEPS_STEP_POSITIVE_COST if positive_cost else 0)
    p = _params(g, t_hop, m, tau, D_max, chi, n_als, als_tol, flags)
    n = len(states)
    hs = (_vp * n)(*[gp.h for gp in states])
    st = np.zeros(n, np.int32)
    rc = lib().peps_batch_trotter_step(hs, n, kind, C.byref(p), st.ctypes.data_as(_i32p))
    if rc:
        bad = [i for i in range(n) if st[i]]
        i = bad[0] if bad else 0
        raise PepsError(rc, f"state {i}: " + (lib().peps_last_error(states[i].ctx.h) or b"").decode())
    return st


def tau_stages(n_ite):
    """The tau annealing of the ite_ground_state drivers: [(tau, steps), ...]."""
    s1, s2, s3, s4, s5 = n_ite // 8, n_ite // 8, n_ite // 4, n_ite // 6, n_ite // 6
    return [(0.1, s1), (0.05, s2), (0.02, s3), (0.01, s4), (0.005, s5), (0.001, n_ite - (s1 + s2 + s3 + s4 + s5))]


def init_gauge_peps(nx, ny, gd, g_charges, noise=0.0, seed=None, ctx=None) -> GaugePEPS:
    """init_gauge_peps (finite_peps_gauge_invariant.jl:180-218) on the device."""
    from . import lattice
    a = lattice.init_gauge_peps_arrays(nx, ny, gd, g_charges, noise, seed)
    return GaugePEPS.from_arrays(a["tensors"], a["lh"], a["lv"], a["qh"], a["qv"], a["gch"], gd,
                                 is_complex=noise > 0, ctx=ctx)


def init_finite_peps_v2(nx, ny, gd, noise=0.0, seed=None, ctx=None) -> GaugePEPS:
    """init_finite_peps_v2 (finite_peps_full_update.jl:29-62) on the device."""
    from . import lattice
    a = lattice.init_finite_peps_v2_arrays(nx, ny, gd, noise, seed)
    return GaugePEPS.from_arrays(a["tensors"], a["lh"], a["lv"], a["qh"], a["qv"], a["gch"], gd,
                                 is_complex=noise > 0, ctx=ctx)


def ite_ground_state(gp: GaugePEPS, kind, g, t_hop, m, D_max, n_ite, chi=64, n_als=30, als_tol=0.0,
                     renorm=False, positive_cost=False) -> int:
    """Run the tau-annealed imaginary-time schedule of the ite_ground_state
    drivers on an existing state in place (kind PEPS_STEP_V4 / _GAUGE /
    _FULL_V2).  Returns the steps completed."""
    flags = (PEPS_STEP_RENORM if renorm else 0) | (PEPS_STEP_POSITIVE_COST if positive_cost else 0)
    p = _params(g, t_hop, m, 0.0, D_max, chi, n_als, als_tol, flags)
    done = C.c_int64()
    gp.ctx.check(lib().peps_ite_ground_state(gp.h, kind, C.byref(p), int(n_ite), C.byref(done)))
    return int(done.value)


def ite_ground_state_v4(nx, ny, gd, D_max, g, t_hop, m, g_charges, n_ite=300, noise=0.1, chi=64, n_als=30,
                        seed=None, ctx=None, **opts) -> GaugePEPS:
    """ite_ground_state_v4 (finite_peps_fullupdate_v4.jl:440-467)."""
    gp = init_gauge_peps(nx, ny, gd, g_charges, noise, seed, ctx)
    ite_ground_state(gp, PEPS_STEP_V4, g, t_hop, m, D_max, n_ite, chi, n_als, **opts)
    return gp


def ite_ground_state_v3(nx, ny, gd, D_max, g, t_hop, m, g_charges, n_ite=600, noise=0.1, seed=None,
                        ctx=None) -> GaugePEPS:
    """ite_ground_state_v3 (finite_peps_gauge_invariant.jl:548-577)."""
    gp = init_gauge_peps(nx, ny, gd, g_charges, noise, seed, ctx)
    ite_ground_state(gp, PEPS_STEP_GAUGE, g, t_hop, m, D_max, n_ite)
    return gp


def ite_ground_state_v2(nx, ny, gd, D_max, g, t_hop, m, n_ite=600, noise=0.01, seed=None,
                        ctx=None) -> GaugePEPS:
    """ite_ground_state_v2 (finite_peps_full_update.jl:582-629, use_gauss=false)."""
    gp = init_finite_peps_v2(nx, ny, gd, noise, seed, ctx)
    ite_ground_state(gp, PEPS_STEP_FULL_V2, g, t_hop, m, D_max, n_ite)
    return gp


def compute_energy_v4(gp: GaugePEPS, g, t_hop, m, chi=64, renorm=False) -> float:
    """compute_energy_v4 (finite_peps_fullupdate_v4.jl:390): total energy
    (renorm: normalised boundaries, as in trotter_step_v4)."""
    p = _params(g, t_hop, m, 0.0, 1, chi, 0, 0.0, PEPS_STEP_RENORM if renorm else 0)
    E = C.c_double()
    gp.ctx.check(lib().peps_energy_v4(gp.h, C.byref(p), C.byref(E)))
    return E.value


def gemm(A, B, a_kfast=False, conjA=False, b_kfast=True, conjB=False, ctx=None):
    """The engine's contraction GEMM: returns op(A) @ op(B) as an (M, N) array.

    A is given as the stored matrix: (M, K) if not a_kfast, (K, M) if a_kfast;
    B as (K, N) if b_kfast, (N, K) if not.
    """
    ctx = ctx or default_context()
    cp = np.iscomplexobj(A) or np.iscomplexobj(B)
    a = _as_array(A, cp)
    b = _as_array(B, cp)
    M, K = (a.shape[1], a.shape[0]) if a_kfast else a.shape
    K2, N = b.shape if b_kfast else (b.shape[1], b.shape[0])
    assert K == K2
    out = np.empty((M, N), dtype=a.dtype, order="F")
    ctx.check(lib().peps_gemm(ctx.h, int(cp), M, N, K, _dptr(a), int(a_kfast), int(conjA), _dptr(b),
                              int(b_kfast), int(conjB), _dptr(out)))
    return out


def gemm_noalloc(M, N, K, a_ptr, b_ptr, c_ptr, is_complex=False, a_kfast=False, conjA=False, b_kfast=True,
                 conjB=False, ctx=None):
    """peps_gemm_noalloc: C = op(A) op(B) on DEVICE buffers the caller owns
    (integer device addresses, column-major, the layouts of gemm), enqueued on
    the context's stream without any allocation; ctx.sync() waits for it."""
    ctx = ctx or default_context()
    ctx.check(lib().peps_gemm_noalloc(ctx.h, int(bool(is_complex)), int(M), int(N), int(K), C.c_void_p(a_ptr),
                                      int(a_kfast), int(conjA), C.c_void_p(b_ptr), int(b_kfast), int(conjB),
                                      C.c_void_p(c_ptr)))


def permute(T, perm, ctx=None):
    """The engine's permutation kernel: numpy.transpose(T, perm) (column-major storage)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(T)
    a = _as_array(T, cp)
    d = np.array(a.shape, dtype=np.int64)
    pr = np.array(perm, dtype=np.int32)
    out = np.empty(tuple(a.shape[i] for i in perm), dtype=a.dtype, order="F")
    ctx.check(lib().peps_permute(ctx.h, int(cp), a.ndim, _iptr(d), pr.ctypes.data_as(_i32p), _dptr(a), _dptr(out)))
    return out


def gemm_acc(A, B, Cin, beta, a_kfast=False, conjA=False, b_kfast=True, conjB=False, ctx=None):
    """op(A) @ op(B) + beta * Cin through the engine's GEMM (layouts as gemm)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(A) or np.iscomplexobj(B) or np.iscomplexobj(Cin)
    a = _as_array(A, cp)
    b = _as_array(B, cp)
    M, K = (a.shape[1], a.shape[0]) if a_kfast else a.shape
    K2, N = b.shape if b_kfast else (b.shape[1], b.shape[0])
    assert K == K2 and Cin.shape == (M, N)
    out = _as_array(Cin, cp).copy(order="F")
    ctx.check(lib().peps_gemm_acc(ctx.h, int(cp), M, N, K, _dptr(a), int(a_kfast), int(conjA), _dptr(b),
                                  int(b_kfast), int(conjB), beta, _dptr(out)))
    return out


def eigh(A, ctx=None):
    """Batched Hermitian eigendecomposition (lower triangles read): A (..., n, n) -> (w, V), w ascending."""
    ctx = ctx or default_context()
    A = np.asarray(A)
    cp = np.iscomplexobj(A)
    n = A.shape[-1]
    batch = A.reshape(-1, n, n)
    nb = batch.shape[0]
    # contiguous column-major matrices, one after the other
    a = np.ascontiguousarray(np.transpose(batch, (0, 2, 1)), dtype=np.complex128 if cp else np.float64)
    V = np.empty_like(a)
    w = np.empty((nb, n))
    ctx.check(lib().peps_eigh(ctx.h, int(cp), n, nb, _dptr(a), _dptr(V), _dptr(w)))
    V = np.transpose(V, (0, 2, 1))
    return w.reshape(A.shape[:-1]), V.reshape(A.shape)


def solve_herm(N, S, reg=1e-10, ctx=None):
    """_solve_herm (finite_peps_fullupdate_v4.jl:37-45): the eigenvalue-floored
    solve of the ALS, for the columns of S (n x k)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(N) or np.iscomplexobj(S)
    a = _as_array(N, cp)
    s2 = np.asarray(S)
    s2 = s2.reshape(-1, 1) if s2.ndim == 1 else s2
    b = _as_array(s2, cp)
    n, k = a.shape[0], b.shape[1]
    assert a.shape == (n, n) and b.shape[0] == n
    out = np.empty((n, k), dtype=a.dtype, order="F")
    ctx.check(lib().peps_solve_herm(ctx.h, int(cp), n, k, _dptr(a), _dptr(b), float(reg), _dptr(out)))
    return out


def svd(A, ctx=None):
    """Thin SVD through the library (QR + one-sided Jacobi): (U, S, Vt)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(A)
    a = _as_array(A, cp)
    m, n = a.shape
    k = min(m, n)
    U = np.empty((m, k), dtype=a.dtype, order="F")
    Vt = np.empty((k, n), dtype=a.dtype, order="F")
    S = np.empty(k)
    ctx.check(lib().peps_svd(ctx.h, int(cp), m, n, _dptr(a), _dptr(U), _dptr(S), _dptr(Vt)))
    return U, S, Vt


def qr(A, ctx=None):
    """Thin QR through the library: (Q, R)."""
    ctx = ctx or default_context()
    cp = np.iscomplexobj(A)
    a = _as_array(A, cp)
    m, n = a.shape
    k = min(m, n)
    Q = np.empty((m, k), dtype=a.dtype, order="F")
    R = np.empty((k, n), dtype=a.dtype, order="F")
    ctx.check(lib().peps_qr(ctx.h, int(cp), m, n, _dptr(a), _dptr(Q), _dptr(R)))
    return Q, R


# ── u1_lgt_hamiltonian.jl ────────────────────────────────────────────────────

def bond_operator(orient, ix, iy, nx, ny, gd, g, t_hop, m, tau=0.0):
    """H_merged_{h,v}_site (tau == 0) or exp(-tau H) (u1_lgt_hamiltonian.jl:829,851,396)."""
    n = C.c_int32()
    L = lib()
    st = L.peps_bond_operator(orient, ix, iy, nx, ny, gd, g, t_hop, m, tau, C.byref(n), None)
    if st != 0:
        raise PepsError(st, "peps_bond_operator failed")
    out = np.zeros((n.value, n.value), order="F")
    st = L.peps_bond_operator(orient, ix, iy, nx, ny, gd, g, t_hop, m, tau, C.byref(n), _dptr(out))
    if st != 0:
        raise PepsError(st, "peps_bond_operator failed")
    return out


# ── finite_ed.jl: exact diagonalisation in one gauge sector ──────────────────

class ExactDiag:
    """One Gauss-law sector of the finite U(1) LGT on the GPU (finite_ed.jl):
    the basis (build_basis), the Hamiltonian (build_hamiltonian), its ground
    state (find_ground_state) and the observables (measure_ED).
    g_charges[ix-1][iy-1] is the background charge of site (ix, iy)."""

    def __init__(self, nx, ny, dg, g_charges, ctx=None):
        self.ctx = ctx or default_context()
        self.nx, self.ny, self.dg = nx, ny, dg
        g = np.asarray(g_charges, dtype=np.int64).reshape(nx, ny)
        gf = np.ascontiguousarray(g.T.reshape(-1))  # ix fastest
        self.h = _vp()
        self.ctx.check(lib().peps_ed_create(self.ctx.h, nx, ny, dg, _iptr(gf), C.byref(self.h)))
        d = C.c_int64()
        self.ctx.check(lib().peps_ed_dim(self.h, C.byref(d)))
        self.dim = d.value

    def __del__(self):
        try:
            if self.h:
                lib().peps_ed_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def keys(self):
        out = np.empty(self.dim, np.uint64)
        self.ctx.check(lib().peps_ed_keys(self.h, out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def hamiltonian(self, g, t, m):
        self.ctx.check(lib().peps_ed_hamiltonian(self.h, float(g), float(t), float(m)))
        return self

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.dim)
        self.ctx.check(lib().peps_ed_apply(self.h, _dptr(x), _dptr(y)))
        return y

    def ground_state(self, tol=1e-10):
        e = C.c_double()
        psi = np.empty(self.dim)
        self.ctx.check(lib().peps_ed_ground_state(self.h, float(tol), C.byref(e), _dptr(psi)))
        return e.value, psi

    def measure(self, psi=None):
        out = np.zeros(4)
        nf = np.zeros(self.nx * self.ny)
        e2 = np.zeros(self.nx * self.ny)
        p = None if psi is None else np.ascontiguousarray(psi, dtype=np.float64)
        self.ctx.check(lib().peps_ed_measure(self.h, None if p is None else _dptr(p), _dptr(out), _dptr(nf),
                                             _dptr(e2)))
        return dict(nf_mean=out[0], nf_even=out[1], nf_odd=out[2], E2_mean=out[3],
                    nf_grid=nf.reshape(self.ny, self.nx).T, E2_grid=e2.reshape(self.ny, self.nx).T)


    def evolve(self, psi0, dt, n_steps, krylovdim=30, tol=1e-10):
        """evolve_ED! (finite_ed.jl:405-430) on the GPU: returns (list of
        measure dicts at t = 0, dt, ..., the final complex state, Krylov
        projections used)."""
        p = np.ascontiguousarray(psi0, dtype=np.complex128)
        if p.shape != (self.dim,):
            raise ValueError("evolve: state of the basis dimension expected")
        ns = self.nx * self.ny
        obs = np.zeros((n_steps + 1, 4))
        nf = np.zeros((n_steps + 1, ns))
        e2 = np.zeros((n_steps + 1, ns))
        out = np.empty(self.dim, np.complex128)
        proj = C.c_int32()
        self.ctx.check(lib().peps_ed_evolve(self.h, p.ctypes.data_as(_dp), float(dt), int(n_steps),
                                            int(min(krylovdim, self.dim)), float(tol), out.ctypes.data_as(_dp),
                                            _dptr(obs), _dptr(nf), _dptr(e2), C.byref(proj)))
        data = [dict(t=s * dt, nf_mean=obs[s, 0], nf_even=obs[s, 1], nf_odd=obs[s, 2], E2_mean=obs[s, 3],
                     nf_grid=nf[s].reshape(self.ny, self.nx).T, E2_grid=e2[s].reshape(self.ny, self.nx).T)
                for s in range(n_steps + 1)]
        return data, out, proj.value


def build_basis(nx, ny, dg, g_charges, ctx=None) -> ExactDiag:
    """build_basis (finite_ed.jl:125)."""
    return ExactDiag(nx, ny, dg, g_charges, ctx=ctx)


def build_hamiltonian(ed: ExactDiag, g, t, m) -> ExactDiag:
    """build_hamiltonian (finite_ed.jl:252), kept on the device."""
    return ed.hamiltonian(g, t, m)


def find_ground_state(ed: ExactDiag, tol=1e-10):
    """find_ground_state (finite_ed.jl:333) -> (E0, psi0)."""
    return ed.ground_state(tol)


def measure_ED(ed: ExactDiag, psi=None):
    """measure_ED (finite_ed.jl:369) -> dict(nf_mean, nf_even, nf_odd, E2_mean, nf_grid, E2_grid)."""
    return ed.measure(psi)
