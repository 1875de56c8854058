"""Python binding (ctypes) of the MI355X DQMC engine's C-ABI (include/dqmc_amd.h).

The engine itself is native: HIP kernels for gfx950 plus a C++ host runtime
in lib/libdqmc_amd.so.  This module only marshals arguments.  It fails
loudly when the library has not been built; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdqmc_amd.so")

OK, ERR_ARG, ERR_HIP, ERR_MEMORY, ERR_STATE, ERR_NUMERIC = 0, -1, -2, -3, -4, -5
REPULSIVE, ATTRACTIVE = 0, 1
PREC_FP64, PREC_MIXED, PREC_FP32_WRAP, PREC_FP32_PRODUCT = 0, 1, 2, 3
STATUS_NONFINITE, STATUS_SINGULAR, STATUS_WRAP_LIMIT, STATUS_PAIR_SYNC = 1, 2, 4, 8
STATUS_FATAL = STATUS_NONFINITE | STATUS_SINGULAR | STATUS_PAIR_SYNC
LAUNCH_AUTO, LAUNCH_SINGLE, LAUNCH_PAIR, LAUNCH_QUAD = 0, 1, 2, 3
DEFAULT_WRAP_LIMIT = 1e-3
SWEEP_MEASURE, SWEEP_RECORD, SWEEP_FROZEN, SWEEP_PROFILE, SWEEP_DYNAMIC = 1, 2, 4, 8, 16
PHASES = ["wrap", "metropolis", "product", "push", "fresh", "measure", "walk", "panels"]
OBS_NAMES = ["sign", "filling_number", "double_occupancy", "kinetic_energy", "momentum_distribution",
             "local_spin_corr", "spin_density_structure_factor", "charge_density_structure_factor",
             "s_wave_pairing_corr"]

EXPORTS = ["dqmc_context_create", "dqmc_context_destroy", "dqmc_context_error", "dqmc_batch_footprint",
           "dqmc_batch_create", "dqmc_batch_destroy", "dqmc_batch_size", "dqmc_batch_use_graph",
           "dqmc_batch_set_launch_mode", "dqmc_batch_last_launch_mode", "dqmc_autotune_launch", "dqmc_fields_random", "dqmc_set_fields",
           "dqmc_get_fields", "dqmc_set_rng_state", "dqmc_set_expk", "dqmc_get_expk", "dqmc_init", "dqmc_sweep",
           "dqmc_set_wrap_error_limit",
           "dqmc_synchronize", "dqmc_last_sweep_ms", "dqmc_get_green", "dqmc_get_slice_green", "dqmc_get_slice_signs",
           "dqmc_get_accepts", "dqmc_get_info", "dqmc_get_measure", "dqmc_measure_recorded", "dqmc_clear_measure",
           "dqmc_get_dynamic_green", "dqmc_get_dynamic_measure", "dqmc_num_kstars", "dqmc_write_fields_file",
           "dqmc_read_fields_file", "dqmc_write_kstars_file", "dqmc_write_tgrids_file", "dqmc_write_observable_file",
           "dqmc_write_observable_bins_file", "dqmc_k_gemm", "dqmc_k_wrap", "dqmc_k_rank1",
           "dqmc_k_udv_push", "dqmc_k_greens_recompute", "dqmc_autotune_delay",
           "dqmc_server_create", "dqmc_server_submit", "dqmc_server_cancel", "dqmc_server_poll",
           "dqmc_server_wait_idle", "dqmc_server_get_stats", "dqmc_server_error", "dqmc_server_destroy",
           "dqmc_server_get_dynamic", "dqmc_workspace_create", "dqmc_workspace_destroy", "dqmc_workspace_expk",
           "dqmc_kd_wrap", "dqmc_kd_update_ratio", "dqmc_kd_rank1", "dqmc_kd_flip_field", "dqmc_kd_udv_push",
           "dqmc_kd_greens_recompute", "dqmc_chain_update_ratio", "dqmc_chain_update_green", "dqmc_chain_flip_field",
           "dqmc_chain_mult_green", "dqmc_set_green", "dqmc_chain_device_buffers",
           "dqmc_k_greens_recompute_noalloc"]


class ChainParams(C.Structure):
    _fields_ = [("model", C.c_int), ("side", C.c_int), ("time_size", C.c_int), ("stabilization_pace", C.c_int),
                ("checkerboard", C.c_int), ("delay", C.c_int), ("momentum_index", C.c_int),
                ("precision", C.c_int), ("beta", C.c_double), ("hopping_t", C.c_double),
                ("onsite_u", C.c_double), ("chemical_potential", C.c_double), ("seed", C.c_longlong)]


class ChainInfo(C.Structure):
    _fields_ = [("sign", C.c_double), ("wrap_error", C.c_double), ("rng", C.c_ulonglong), ("pos", C.c_int),
                ("status", C.c_int), ("accepted", C.c_longlong), ("sweeps", C.c_longlong),
                ("phase_seconds", C.c_double * 8), ("paired", C.c_int), ("kron", C.c_int)]


JOB_DYNAMIC = 1


class Job(C.Structure):
    _fields_ = [("chain", ChainParams), ("warmup_sweeps", C.c_int), ("measure_sweeps", C.c_int),
                ("bin_size", C.c_int), ("options", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_longlong), ("status", C.c_int), ("chain_status", C.c_int), ("n_bins", C.c_int),
                ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_finish", C.c_double), ("sweeps", C.c_longlong),
                ("wrap_error", C.c_double), ("mean", C.c_double * 9), ("error", C.c_double * 9)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("active", C.c_longlong),
                ("queued", C.c_longlong), ("chain_sweeps", C.c_longlong), ("launches", C.c_longlong),
                ("bytes_in_use", C.c_size_t), ("bytes_peak", C.c_size_t), ("budget", C.c_size_t),
                ("elapsed_s", C.c_double)]


class DqmcError(RuntimeError):
    pass


_lib = None


def lib():
    """Load lib/libdqmc_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DqmcError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        vp, ip = C.c_void_p, C.c_int
        L.dqmc_context_create.argtypes = [ip, C.POINTER(vp)]
        L.dqmc_context_destroy.argtypes = [vp]
        L.dqmc_context_error.argtypes = [vp]
        L.dqmc_context_error.restype = C.c_char_p
        L.dqmc_batch_footprint.argtypes = [ip, C.POINTER(ChainParams), ip, C.POINTER(C.c_size_t)]
        L.dqmc_batch_create.argtypes = [vp, ip, C.POINTER(ChainParams), ip, C.POINTER(vp)]
        for n in ["dqmc_batch_destroy", "dqmc_batch_size", "dqmc_fields_random", "dqmc_init", "dqmc_synchronize",
                  "dqmc_clear_measure", "dqmc_measure_recorded"]:
            getattr(L, n).argtypes = [vp]
        L.dqmc_batch_use_graph.argtypes = [vp, ip]
        L.dqmc_batch_set_launch_mode.argtypes = [vp, ip]
        L.dqmc_batch_last_launch_mode.argtypes = [vp, C.POINTER(C.c_int)]
        L.dqmc_set_fields.argtypes = [vp, ip, vp]
        L.dqmc_get_fields.argtypes = [vp, ip, vp]
        L.dqmc_set_rng_state.argtypes = [vp, ip, C.c_ulonglong]
        L.dqmc_set_expk.argtypes = [vp, ip, vp, vp]
        L.dqmc_get_expk.argtypes = [vp, ip, vp, vp]
        L.dqmc_sweep.argtypes = [vp, ip, ip]
        L.dqmc_set_wrap_error_limit.argtypes = [vp, C.c_double]
        L.dqmc_last_sweep_ms.argtypes = [vp, C.POINTER(C.c_float)]
        L.dqmc_get_green.argtypes = [vp, ip, ip, vp]
        L.dqmc_get_slice_green.argtypes = [vp, ip, ip, ip, vp]
        L.dqmc_get_slice_signs.argtypes = [vp, ip, vp]
        L.dqmc_get_accepts.argtypes = [vp, ip, vp]
        L.dqmc_get_info.argtypes = [vp, ip, C.POINTER(ChainInfo)]
        L.dqmc_get_measure.argtypes = [vp, ip, vp, vp]
        L.dqmc_get_dynamic_green.argtypes = [vp, ip, ip, ip, ip, vp]
        L.dqmc_get_dynamic_measure.argtypes = [vp, ip, ip, vp, vp, vp, vp, vp]
        L.dqmc_num_kstars.argtypes = [vp, ip, vp]
        cp = C.c_char_p
        L.dqmc_write_fields_file.argtypes = [cp, ip, ip, vp]
        L.dqmc_read_fields_file.argtypes = [cp, ip, ip, vp]
        L.dqmc_write_kstars_file.argtypes = [cp, ip]
        L.dqmc_write_tgrids_file.argtypes = [cp, ip, C.c_double]
        L.dqmc_write_observable_file.argtypes = [cp, ip, ip, ip, vp, vp]
        L.dqmc_write_observable_bins_file.argtypes = [cp, ip, ip, ip, ip, vp]
        L.dqmc_k_gemm.argtypes = [vp, ip, ip, ip, ip, ip, ip, vp, vp, vp, C.c_double, vp, vp, vp]
        L.dqmc_k_wrap.argtypes = [vp, C.POINTER(ChainParams), vp, vp, vp, ip, ip, ip, vp]
        L.dqmc_k_rank1.argtypes = [vp, C.POINTER(ChainParams), vp, ip, ip, vp, vp]
        L.dqmc_k_udv_push.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp, vp]
        L.dqmc_k_greens_recompute.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp, vp]
        L.dqmc_autotune_delay.argtypes = [vp, C.POINTER(ChainParams), ip, vp, ip, ip, C.POINTER(ip), vp]
        L.dqmc_autotune_launch.argtypes = [vp, C.POINTER(ChainParams), ip, ip, C.POINTER(ip), vp]
        L.dqmc_server_create.argtypes = [vp, C.c_size_t, ip, ip, C.POINTER(vp)]
        L.dqmc_server_submit.argtypes = [vp, C.POINTER(Job), C.POINTER(C.c_longlong)]
        L.dqmc_server_cancel.argtypes = [vp, C.c_longlong]
        L.dqmc_server_poll.argtypes = [vp, C.POINTER(JobResult), ip, C.POINTER(ip)]
        L.dqmc_server_wait_idle.argtypes = [vp, C.c_double]
        L.dqmc_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.dqmc_server_error.argtypes = [vp]
        L.dqmc_server_error.restype = C.c_char_p
        L.dqmc_server_destroy.argtypes = [vp]
        L.dqmc_server_get_dynamic.argtypes = [vp, C.c_longlong, C.POINTER(ip), C.POINTER(ip), vp, vp]
        L.dqmc_workspace_create.argtypes = [vp, C.POINTER(ChainParams), C.POINTER(vp)]
        L.dqmc_workspace_destroy.argtypes = [vp]
        L.dqmc_workspace_expk.argtypes = [vp, vp, vp]
        L.dqmc_kd_wrap.argtypes = [vp, vp, vp, vp, ip, ip, ip, vp, vp]
        L.dqmc_kd_update_ratio.argtypes = [vp, vp, ip, ip, vp, vp, vp, vp]
        L.dqmc_kd_rank1.argtypes = [vp, vp, ip, ip, vp, vp, vp]
        L.dqmc_kd_flip_field.argtypes = [vp, vp, ip, ip, vp]
        L.dqmc_kd_udv_push.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.dqmc_kd_greens_recompute.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.dqmc_chain_update_ratio.argtypes = [vp, ip, ip, ip, C.POINTER(C.c_double)]
        L.dqmc_chain_update_green.argtypes = [vp, ip, ip, ip]
        L.dqmc_chain_flip_field.argtypes = [vp, ip, ip, ip]
        L.dqmc_chain_mult_green.argtypes = [vp, ip, ip, ip, ip]
        L.dqmc_set_green.argtypes = [vp, ip, ip, vp]
        L.dqmc_chain_device_buffers.argtypes = [vp, ip, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
        L.dqmc_k_greens_recompute_noalloc.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_params(side, time_size, beta, pace=10, model=REPULSIVE, checkerboard=False, delay=16, momentum=-1,
                t=1.0, U=4.0, mu=0.0, seed=42, precision=PREC_FP64):
    """One chain's parameters; attractive chains take |U|."""
    return ChainParams(int(model), int(side), int(time_size), int(pace), int(bool(checkerboard)), int(delay),
                       int(momentum), int(precision), float(beta), float(t), float(U), float(mu), int(seed))


def footprint(params, record=False):
    arr = (ChainParams * len(params))(*params)
    out = C.c_size_t(0)
    rc = lib().dqmc_batch_footprint(len(params), arr, int(record), C.byref(out))
    if rc != OK:
        raise DqmcError(f"dqmc_batch_footprint failed ({rc})")
    return out.value


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().dqmc_context_create(device, C.byref(self.h))
        if rc != OK:
            raise DqmcError(f"dqmc_context_create({device}) failed ({rc})")

    def gemm(self, variant, A, B, C=None, transA=False, transB=False, beta=0.0, rs=None, ms=None, cs=None):
        """Run one workgroup GEMM on the device (test entry point)."""
        A = np.asfortranarray(A, dtype=np.float64); B = np.asfortranarray(B, dtype=np.float64)
        M = A.shape[1] if transA else A.shape[0]
        K = A.shape[0] if transA else A.shape[1]
        Nc = B.shape[0] if transB else B.shape[1]
        C = np.zeros((M, Nc), order="F") if C is None else np.array(C, dtype=np.float64, order="F", copy=True)
        vec = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (rs, ms, cs)]
        ptr = [None if v is None else _p(v) for v in vec]
        self.check(lib().dqmc_k_gemm(self.h, variant, M, Nc, K, int(transA), int(transB), _p(A), _p(B), _p(C),
                                     float(beta), ptr[0], ptr[1], ptr[2]), "dqmc_k_gemm")
        return C

    # B-matrix products by name: ModelBase mult_* (src/model/repulsive_hubbard.cpp:158-245) and the wraps
    WRAP_OPS = {"B_left": 0, "B_right": 1, "invB_left": 2, "invB_right": 3, "transB_left": 4,
                "wrap_up": 5, "wrap_down": 6}

    def k_wrap(self, params, fields, t, spin, op, M, expK=None, invExpK=None):
        """One slice's B-matrix product (or wrap) applied to M on the device; returns the result.
        t is the reference's time index (field row t - 1, t = 0 the slice at beta)."""
        op = self.WRAP_OPS.get(op, op)
        f = np.ascontiguousarray(fields, dtype=np.int8)
        M = np.array(M, dtype=np.float64, order="F", copy=True)
        ek = None if expK is None else np.asfortranarray(expK, dtype=np.float64)
        ik = None if invExpK is None else np.asfortranarray(invExpK, dtype=np.float64)
        self.check(lib().dqmc_k_wrap(self.h, C.byref(params), _p(f), None if ek is None else _p(ek),
                                     None if ik is None else _p(ik), int(t), int(spin), int(op), _p(M)),
                   "dqmc_k_wrap")
        return M

    def k_rank1(self, params, fields, t, i, G_up, G_dn=None):
        """Single-flip update of the Green's functions for the field at (t, i); returns (G_up, G_dn)."""
        f = np.ascontiguousarray(fields, dtype=np.int8)
        gu = np.array(G_up, dtype=np.float64, order="F", copy=True)
        gd = None if G_dn is None else np.array(G_dn, dtype=np.float64, order="F", copy=True)
        self.check(lib().dqmc_k_rank1(self.h, C.byref(params), _p(f), int(t), int(i), _p(gu),
                                      None if gd is None else _p(gd)), "dqmc_k_rank1")
        return gu, gd

    def k_udv_push(self, A, top=None):
        """Push A onto a stack whose top is (U, D, V) (None: empty); returns the new (U, D, V)
        with U diag(D) V^T = A U_top diag(D_top) V_top^T."""
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        U = np.zeros((n, n), order="F"); D = np.zeros(n); V = np.zeros((n, n), order="F")
        t = [None] * 3 if top is None else [np.asfortranarray(top[0], dtype=np.float64),
                                             np.ascontiguousarray(top[1], dtype=np.float64),
                                             np.asfortranarray(top[2], dtype=np.float64)]
        self.check(lib().dqmc_k_udv_push(self.h, n, _p(A), *[None if x is None else _p(x) for x in t],
                                         _p(U), _p(D), _p(V)), "dqmc_k_udv_push")
        return U, D, V

    def k_greens_recompute(self, left=None, right=None):
        """(1 + B_L B_R^T)^-1 from the stack tops left = (U, D, V), right = (U, D, V); None: empty."""
        side = left if left is not None else right
        n = np.asarray(side[0]).shape[0]
        args = []
        for top in (left, right):
            if top is None:
                args += [None] * 3
            else:
                args += [np.asfortranarray(top[0], dtype=np.float64), np.ascontiguousarray(top[1], dtype=np.float64),
                         np.asfortranarray(top[2], dtype=np.float64)]
        G = np.zeros((n, n), order="F")
        self.check(lib().dqmc_k_greens_recompute(self.h, n, *[None if x is None else _p(x) for x in args], _p(G)),
                   "dqmc_k_greens_recompute")
        return G

    def k_greens_recompute_noalloc(self, N, left, right, G, status=None, stream=None):
        """The recompute on device buffers (addresses) with the library's static scratch."""
        ul, dl, vl = left if left is not None else (None, None, None)
        ur, dr, vr = right if right is not None else (None, None, None)
        self.check(lib().dqmc_k_greens_recompute_noalloc(self.h, N, ul, dl, vl, ur, dr, vr, G, status, stream),
                   "dqmc_k_greens_recompute_noalloc")

    def autotune_delay(self, params, n_chains, candidates=(8, 16, 32), sweeps=2):
        """Fastest delayed-update depth for chains of this shape (times every candidate)."""
        cand = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(cand), dtype=np.float32)
        best = C.c_int(0)
        self.check(lib().dqmc_autotune_delay(self.h, C.byref(params), int(n_chains), _p(cand), len(cand), int(sweeps),
                                             C.byref(best), _p(ms)), "dqmc_autotune_delay")
        return best.value, dict(zip([int(c) for c in cand], [float(x) for x in ms]))

    def autotune_launch(self, params, n_chains, sweeps=2):
        """Faster chain -> workgroup mapping (LAUNCH_SINGLE or LAUNCH_PAIR) for chains of this shape, and the
        launch times ({mode: ms}; pairs are left out when the batch does not qualify)."""
        ms = np.full(2, -1.0, dtype=np.float32)
        best = C.c_int(0)
        self.check(lib().dqmc_autotune_launch(self.h, C.byref(params), int(n_chains), int(sweeps), C.byref(best), _p(ms)),
                   "dqmc_autotune_launch")
        return best.value, {m: float(t) for m, t in zip((LAUNCH_SINGLE, LAUNCH_PAIR), ms) if t >= 0}

    def error(self):
        return lib().dqmc_context_error(self.h).decode()

    def check(self, rc, what):
        if rc != OK:
            raise DqmcError(f"{what} failed ({rc}): {self.error()}")

    def close(self):
        if self.h:
            lib().dqmc_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Batch:
    """A batch of independent chains resident on one GPU."""

    def __init__(self, ctx, params, record=False):
        self.ctx = ctx
        self.params = list(params)
        self.n = len(self.params)
        arr = (ChainParams * self.n)(*self.params)
        self.h = C.c_void_p()
        ctx.check(lib().dqmc_batch_create(ctx.h, self.n, arr, int(record), C.byref(self.h)), "dqmc_batch_create")
        self.N = [p.side * p.side for p in self.params]
        self.L = [p.time_size for p in self.params]

    def close(self):
        if getattr(self, "h", None):
            lib().dqmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def use_graph(self, enable=True):
        self.ctx.check(lib().dqmc_batch_use_graph(self.h, int(enable)), "dqmc_batch_use_graph")

    def set_launch_mode(self, mode):
        """LAUNCH_AUTO (default), LAUNCH_SINGLE (one workgroup per chain) or LAUNCH_PAIR (one per
        spin of every repulsive chain); raises when the batch does not qualify, and for the
        reserved LAUNCH_QUAD."""
        self.ctx.check(lib().dqmc_batch_set_launch_mode(self.h, int(mode)), "dqmc_batch_set_launch_mode")

    def last_launch_mode(self):
        m = C.c_int(0)
        self.ctx.check(lib().dqmc_batch_last_launch_mode(self.h, C.byref(m)), "dqmc_batch_last_launch_mode")
        return m.value

    def fields_random(self):
        self.ctx.check(lib().dqmc_fields_random(self.h), "dqmc_fields_random")

    def set_fields(self, c, f):
        f = np.ascontiguousarray(f, dtype=np.int8).reshape(self.L[c], self.N[c])
        self.ctx.check(lib().dqmc_set_fields(self.h, c, _p(f)), "dqmc_set_fields")

    def fields(self, c):
        f = np.zeros((self.L[c], self.N[c]), np.int8)
        self.ctx.check(lib().dqmc_get_fields(self.h, c, _p(f)), "dqmc_get_fields")
        return f

    def set_rng_state(self, c, x):
        self.ctx.check(lib().dqmc_set_rng_state(self.h, c, int(x)), "dqmc_set_rng_state")

    def set_expk(self, c, expK, invExpK):
        a = np.asfortranarray(expK, dtype=np.float64); b = np.asfortranarray(invExpK, dtype=np.float64)
        self.ctx.check(lib().dqmc_set_expk(self.h, c, _p(a), _p(b)), "dqmc_set_expk")

    def expk(self, c):
        n = self.N[c]
        a = np.zeros((n, n), order="F"); b = np.zeros((n, n), order="F")
        self.ctx.check(lib().dqmc_get_expk(self.h, c, _p(a), _p(b)), "dqmc_get_expk")
        return a, b

    def init(self):
        self.ctx.check(lib().dqmc_init(self.h), "dqmc_init")

    def sweep(self, n, measure=False, record=False, frozen=False, profile=False, dynamic=False):
        """n single-direction sweeps; with dynamic=True every 0->beta sweep is the time-displaced
        one (DqmcWalker::sweep_for_dynamic_greens) and measures the dynamic observables."""
        flags = ((SWEEP_MEASURE if measure else 0) | (SWEEP_RECORD if record else 0) | (SWEEP_FROZEN if frozen else 0)
                 | (SWEEP_PROFILE if profile else 0) | (SWEEP_DYNAMIC if dynamic else 0))
        self.ctx.check(lib().dqmc_sweep(self.h, int(n), flags), "dqmc_sweep")

    def last_sweep_ms(self):
        ms = C.c_float(0)
        self.ctx.check(lib().dqmc_last_sweep_ms(self.h, C.byref(ms)), "dqmc_last_sweep_ms")
        return ms.value

    def synchronize(self):
        self.ctx.check(lib().dqmc_synchronize(self.h), "dqmc_synchronize")

    def green(self, c, spin):
        n = self.N[c]
        g = np.zeros((n, n), order="F")
        self.ctx.check(lib().dqmc_get_green(self.h, c, spin, _p(g)), "dqmc_get_green")
        return g

    def slice_green(self, c, slot, spin):
        n = self.N[c]
        g = np.zeros((n, n), order="F")
        self.ctx.check(lib().dqmc_get_slice_green(self.h, c, slot, spin, _p(g)), "dqmc_get_slice_green")
        return g

    def slice_signs(self, c):
        s = np.zeros(self.L[c])
        self.ctx.check(lib().dqmc_get_slice_signs(self.h, c, _p(s)), "dqmc_get_slice_signs")
        return s

    def accepts(self, c):
        a = np.zeros((self.L[c], self.N[c]), np.uint8)
        self.ctx.check(lib().dqmc_get_accepts(self.h, c, _p(a)), "dqmc_get_accepts")
        return a

    def info(self, c):
        i = ChainInfo()
        self.ctx.check(lib().dqmc_get_info(self.h, c, C.byref(i)), "dqmc_get_info")
        return i

    def set_wrap_error_limit(self, limit):
        self.ctx.check(lib().dqmc_set_wrap_error_limit(self.h, float(limit)), "dqmc_set_wrap_error_limit")

    # per-site plugin calls of the reference's Metropolis loop on the resident chain
    def update_ratio(self, c, t, i):
        r = C.c_double(0.0)
        self.ctx.check(lib().dqmc_chain_update_ratio(self.h, c, t, i, C.byref(r)), "dqmc_chain_update_ratio")
        return r.value

    def update_green(self, c, t, i):
        self.ctx.check(lib().dqmc_chain_update_green(self.h, c, t, i), "dqmc_chain_update_green")

    def flip_field(self, c, t, i):
        self.ctx.check(lib().dqmc_chain_flip_field(self.h, c, t, i), "dqmc_chain_flip_field")

    def mult_green(self, c, t, spin, op):
        self.ctx.check(lib().dqmc_chain_mult_green(self.h, c, t, spin, op), "dqmc_chain_mult_green")

    def set_green(self, c, spin, g):
        g = np.asfortranarray(g, dtype=np.float64)
        self.ctx.check(lib().dqmc_set_green(self.h, c, spin, _p(g)), "dqmc_set_green")

    def device_buffers(self, c):
        """(G_up, G_dn, fields, stream) device addresses of chain c's resident state."""
        gu, gd, f, st = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        self.ctx.check(lib().dqmc_chain_device_buffers(self.h, c, C.byref(gu), C.byref(gd), C.byref(f), C.byref(st)),
                       "dqmc_chain_device_buffers")
        return gu.value, gd.value, f.value, st.value

    def measure_accum(self, c):
        s = np.zeros(9); n = np.zeros(9)
        self.ctx.check(lib().dqmc_get_measure(self.h, c, _p(s), _p(n)), "dqmc_get_measure")
        return s, n

    def measure_recorded(self):
        self.ctx.check(lib().dqmc_measure_recorded(self.h), "dqmc_measure_recorded")

    def clear_measure(self):
        self.ctx.check(lib().dqmc_clear_measure(self.h), "dqmc_clear_measure")

    def kstars(self, c):
        nk = lib().dqmc_num_kstars(self.h, c, None)
        if nk < 0:
            self.ctx.check(nk, "dqmc_num_kstars")
        k = np.zeros((nk, 2))
        lib().dqmc_num_kstars(self.h, c, _p(k))
        return k

    def dynamic_green(self, c, slot, spin, which):
        """which 0: G(t,0), 1: G(0,t) of slot t-1 kept by the last dynamic RECORD sweep."""
        n = self.N[c]
        g = np.zeros((n, n), order="F")
        self.ctx.check(lib().dqmc_get_dynamic_green(self.h, c, slot, spin, which, _p(g)), "dqmc_get_dynamic_green")
        return g

    def dynamic_accum(self, c, k_index=None):
        """Dynamic sums: gk (n_k x L, all k-stars unless k_index is given), dos (L), dss (L), rho,
        sign, count (MeasureHandler::dynamic_measure, src/measure/measure_methods.cpp:264-461)."""
        L = self.L[c]
        if k_index is None:
            k_index = np.arange(lib().dqmc_num_kstars(self.h, c, None))
        k_index = np.ascontiguousarray(k_index, dtype=np.int32)
        nk = len(k_index)
        gk = np.zeros(nk * L); dos = np.zeros(L); dss = np.zeros(L); sc = np.zeros(3)
        self.ctx.check(lib().dqmc_get_dynamic_measure(self.h, c, nk, _p(k_index), _p(gk), _p(dos), _p(dss), _p(sc)),
                       "dqmc_get_dynamic_measure")
        return dict(gk=gk.reshape(L, nk).T.copy(), dos=dos, dss=dss, rho=sc[0], sign=sc[1], count=sc[2])


# ---------------------------------------------------------------------------
# On-disk formats (QuantumMonteCarlo::DqmcIO, include/dqmc_io.h), written and read by the
# library's host code
# ---------------------------------------------------------------------------
def _io(rc, what, path):
    if rc != 0:
        raise DqmcError(f"{what}({path!r}) failed with status {rc}")


def write_fields(path, fields):
    """DqmcIO::output_bosonic_fields: fields (L x N of +-1) to the reference's text format."""
    f = np.ascontiguousarray(fields, dtype=np.int8)
    _io(lib().dqmc_write_fields_file(os.fsencode(path), f.shape[0], f.shape[1], _p(f)), "dqmc_write_fields_file", path)


def read_fields(path, time_size, space_size):
    """DqmcIO::read_bosonic_fields_from_file for a chain of time_size x space_size fields."""
    f = np.ones((time_size, space_size), np.int8)
    _io(lib().dqmc_read_fields_file(os.fsencode(path), time_size, space_size, _p(f)), "dqmc_read_fields_file", path)
    return f


def write_kstars(path, side):
    _io(lib().dqmc_write_kstars_file(os.fsencode(path), side), "dqmc_write_kstars_file", path)


def write_tgrids(path, time_size, beta):
    _io(lib().dqmc_write_tgrids_file(os.fsencode(path), time_size, float(beta)), "dqmc_write_tgrids_file", path)


def write_observable(path, mean, err):
    """DqmcIO::output_observable (file): a scalar, a vector or a matrix (n_k x L for G(k,t))."""
    m = np.asfortranarray(np.atleast_1d(np.asarray(mean, dtype=np.float64)))
    e = np.asfortranarray(np.atleast_1d(np.asarray(err, dtype=np.float64)))
    rank = np.ndim(mean)
    rows = m.shape[0] if rank >= 1 else 1
    cols = m.shape[1] if rank == 2 else 1
    _io(lib().dqmc_write_observable_file(os.fsencode(path), rank, rows, cols, _p(m), _p(e)),
        "dqmc_write_observable_file", path)


def write_bins(path, bins):
    """DqmcIO::output_observable_in_bins: bins has the bin index first (n_bins, [rows, [cols]])."""
    b = np.asarray(bins, dtype=np.float64)
    rank = b.ndim - 1
    rows = b.shape[1] if rank >= 1 else 1
    cols = b.shape[2] if rank == 2 else 1
    flat = np.ascontiguousarray(np.stack([np.asfortranarray(x).ravel(order="F") for x in b]) if rank == 2 else b)
    _io(lib().dqmc_write_observable_bins_file(os.fsencode(path), rank, b.shape[0], rows, cols, _p(flat)),
        "dqmc_write_observable_bins_file", path)


def normalize_bin(sums, counts):
    """MeasureHandler::normalize_stats (src/measure/measure_handler.cpp:170-189)."""
    sums = np.asarray(sums, dtype=np.float64); counts = np.asarray(counts, dtype=np.float64)
    sign = sums[0] / counts[0]
    out = sums / (counts * sign)
    out[0] = abs(sign)
    return out


def bin_stats(bins):
    """Observable::analyse (include/measure/observable.h:169-193)."""
    b = np.asarray(bins, dtype=np.float64)
    nb = b.shape[0]
    mean = b.mean(axis=0)
    err = np.sqrt(np.maximum((b ** 2).mean(axis=0) - mean ** 2, 0.0)) / np.sqrt(max(nb - 1, 1))
    return mean, err


def result_dict(r):
    return {"id": r.id, "status": r.status, "n_bins": r.n_bins, "t_submit": r.t_submit, "t_start": r.t_start,
            "t_finish": r.t_finish, "latency": r.t_finish - r.t_submit, "sweeps": r.sweeps,
            "wrap_error": r.wrap_error, "chain_status": r.chain_status, "mean": list(r.mean),
            "error": list(r.error)}


class Workspace:
    """Scratch of one chain shape for the building blocks on caller-provided device buffers
    (dqmc_workspace_* and dqmc_kd_* of the C-ABI).  Buffer arguments are device addresses
    (ints, e.g. torch tensors' data_ptr()); stream is a hipStream_t address or None."""

    def __init__(self, ctx, params):
        self.ctx = ctx
        self.h = C.c_void_p()
        self.N = params.side * params.side
        ctx.check(lib().dqmc_workspace_create(ctx.h, C.byref(params), C.byref(self.h)), "dqmc_workspace_create")

    def close(self):
        if self.h:
            lib().dqmc_workspace_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def expk(self):
        a = np.zeros((self.N, self.N), order="F"); b = np.zeros((self.N, self.N), order="F")
        self.ctx.check(lib().dqmc_workspace_expk(self.h, _p(a), _p(b)), "dqmc_workspace_expk")
        return a, b

    def wrap(self, fields, t, spin, op, M, stream=None, expK=None, invExpK=None):
        self.ctx.check(lib().dqmc_kd_wrap(self.h, fields, expK, invExpK, t, spin, op, M, stream), "dqmc_kd_wrap")

    def update_ratio(self, fields, t, i, G_up, G_dn, ratio, stream=None):
        self.ctx.check(lib().dqmc_kd_update_ratio(self.h, fields, t, i, G_up, G_dn, ratio, stream),
                       "dqmc_kd_update_ratio")

    def rank1(self, fields, t, i, G_up, G_dn, stream=None):
        self.ctx.check(lib().dqmc_kd_rank1(self.h, fields, t, i, G_up, G_dn, stream), "dqmc_kd_rank1")

    def flip_field(self, fields, t, i, stream=None):
        self.ctx.check(lib().dqmc_kd_flip_field(self.h, fields, t, i, stream), "dqmc_kd_flip_field")

    def udv_push(self, A, top, U, D, V, status=None, stream=None):
        ui, di, vi = top if top is not None else (None, None, None)
        self.ctx.check(lib().dqmc_kd_udv_push(self.h, A, ui, di, vi, U, D, V, status, stream), "dqmc_kd_udv_push")

    def greens_recompute(self, left, right, G, status=None, stream=None):
        ul, dl, vl = left if left is not None else (None, None, None)
        ur, dr, vr = right if right is not None else (None, None, None)
        self.ctx.check(lib().dqmc_kd_greens_recompute(self.h, ul, dl, vl, ur, dr, vr, G, status, stream),
                       "dqmc_kd_greens_recompute")


class Server:
    """Continuous-batching job server (dqmc_server_* of the C-ABI)."""

    def __init__(self, ctx, hbm_budget_bytes, max_active=4096, lanes=4):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(lib().dqmc_server_create(ctx.h, int(hbm_budget_bytes), int(max_active), int(lanes),
                                           C.byref(self.h)), "dqmc_server_create")

    def submit(self, params, warmup, measure, bin_size, dynamic=False):
        j = Job(params, int(warmup), int(measure), int(bin_size), JOB_DYNAMIC if dynamic else 0)
        jid = C.c_longlong(0)
        rc = lib().dqmc_server_submit(self.h, C.byref(j), C.byref(jid))
        if rc != OK:
            raise DqmcError(f"dqmc_server_submit failed ({rc}): {lib().dqmc_server_error(self.h).decode()}")
        return jid.value

    def cancel(self, jid):
        return lib().dqmc_server_cancel(self.h, int(jid)) == OK

    def dynamic_result(self, jid):
        """Mean and error of the dynamic observables of a finished dynamic job: dicts with
        sign, rho, gk (n_k x L), dos (L), dss (L)."""
        nk, L = C.c_int(0), C.c_int(0)
        rc = lib().dqmc_server_get_dynamic(self.h, int(jid), C.byref(nk), C.byref(L), None, None)
        if rc != OK:
            raise DqmcError(f"dqmc_server_get_dynamic failed ({rc}): {lib().dqmc_server_error(self.h).decode()}")
        n = 2 + nk.value * L.value + 2 * L.value
        m = np.zeros(n); e = np.zeros(n)
        lib().dqmc_server_get_dynamic(self.h, int(jid), C.byref(nk), C.byref(L), _p(m), _p(e))

        def split(v):
            k, t = nk.value, L.value
            return dict(sign=v[0], rho=v[1], gk=v[2:2 + k * t].reshape(t, k).T.copy(),
                        dos=v[2 + k * t:2 + k * t + t].copy(), dss=v[2 + k * t + t:].copy())
        return split(m), split(e)

    def poll(self, max_results=256):
        buf = (JobResult * max_results)()
        n = C.c_int(0)
        rc = lib().dqmc_server_poll(self.h, buf, max_results, C.byref(n))
        if rc != OK:
            raise DqmcError(f"server failure ({rc})")
        return [result_dict(buf[k]) for k in range(n.value)]

    def wait_idle(self, timeout_s):
        return lib().dqmc_server_wait_idle(self.h, float(timeout_s)) == OK

    def stats(self):
        st = ServerStats()
        lib().dqmc_server_get_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().dqmc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
