"""Python view of the MI355X CU-BENs solve path (libcubens_amd.so, include/cubens.h).

The product is the C ABI; this module is a thin ctypes mirror used by tests and
the benchmark.  There is no CPU fallback: every call goes through the HIP
library, and opening a model fails loudly when the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libcubens_amd.so")

SOLVER_BAND64 = 0
SOLVER_BAND32_FGMRES = 1
SOLVER_BJ_FGMRES = 2
SOLVER_LEARNED_FGMRES = 3
SOLVER_BAND32 = 4

GEN_BASE, GEN_I, GEN_IP = 0, 1, 2


class Info(C.Structure):
    _fields_ = [("nj", C.c_long), ("ne_tr", C.c_long), ("ne_fr", C.c_long), ("ne_sh", C.c_long),
                ("neq", C.c_long), ("lss", C.c_long), ("bandwidth", C.c_long),
                ("nnz_blocks", C.c_long), ("bs", C.c_int), ("anaflag", C.c_int),
                ("algflag", C.c_int), ("ncolours_tr", C.c_int), ("ncolours_sh", C.c_int)]


class Options(C.Structure):
    _fields_ = [("solver", C.c_int), ("krylov_tol", C.c_double), ("restart", C.c_int),
                ("maxit", C.c_int), ("judge", C.c_int), ("max_steps", C.c_int),
                ("keep_d", C.c_int), ("kernel_timing", C.c_int), ("factor_schedule", C.c_int),
                ("kernel_timing_every", C.c_int), ("sweep_schedule", C.c_int), ("keep_forces", C.c_int),
                ("checkpoint_path", C.c_char_p), ("checkpoint_exact", C.c_int), ("graph", C.c_int),
                ("refine", C.c_int), ("autotune", C.c_int), ("fp32_elements", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("t_total", C.c_double), ("t_tangent", C.c_double), ("t_factor", C.c_double),
                ("t_solve", C.c_double), ("t_forces", C.c_double), ("t_judge", C.c_double),
                ("t_other", C.c_double), ("newton", C.c_long), ("solves", C.c_long),
                ("factors", C.c_long), ("krylov_iters", C.c_long),
                ("krylov_fallbacks", C.c_long), ("worst_relres", C.c_double),
                ("status", C.c_int), ("nsteps", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


KT_COLUMN, KT_DIAG, KT_PANEL, KT_UPDATE, KT_FWD, KT_BWD, KT_PREP = range(7)
# Per-step driver events (cb_get_step_events): det flag, quad imaginary-root and
# negative-gamma retries, dnorm restarts, itemax halvings.
EV_DET, EV_IMAG, EV_NEG, EV_DNORM, EV_HALVE, EV_YIELD, EV_UNLOAD = range(7)
NEV = 7
KT_NAMES = ("band_column", "band_diag", "band_panel", "band_update", "band_fwd_sweep", "band_bwd_sweep",
            "band_sweep_prep")
FACTOR_FUSED, FACTOR_SPLIT = 0, 1
SWEEP_CHAIN, SWEEP_ROWS, SWEEP_LOOK2, SWEEP_LOOK3 = 0, 1, 2, 3


def options(solver: int = SOLVER_BAND64, krylov_tol: float = 1e-8, restart: int = 30,
            maxit: int = 20000, judge: int = 1, max_steps: int = 1 << 30, keep_d: int = 1,
            kernel_timing: int = 0, factor_schedule: int = 0, kernel_timing_every: int = 1,
            sweep_schedule: int = 0, keep_forces: int = 0, checkpoint_path: str | None = None,
            checkpoint_exact: int = 0, graph: int = 0, refine: int = 0, autotune: int = 0,
            fp32_elements: int = 0) -> Options:
    """cb_options; checkpoint_path is the Newmark checkpoint file (written every
    CHKPT time steps, read first when the model has RFLAG 1); refine the
    iterative refinement steps of the FP32 precision mode's solves."""
    return Options(solver, krylov_tol, restart, maxit, judge, max_steps, keep_d, kernel_timing,
                   factor_schedule, kernel_timing_every, sweep_schedule, keep_forces,
                   checkpoint_path.encode() if checkpoint_path else None, checkpoint_exact, graph, refine,
                   autotune, fp32_elements)


_lib = None

# Every entry point of include/cubens.h with its ctypes signature.
PD = C.POINTER(C.c_double)
PL = C.POINTER(C.c_long)
PI = C.POINTER(C.c_int)
PI32 = C.POINTER(C.c_int32)
PF = C.POINTER(C.c_float)


class BatchStats(C.Structure):
    _fields_ = [("t_total", C.c_double), ("rounds", C.c_long), ("newton", C.c_long),
                ("steps", C.c_long), ("failed", C.c_int), ("ms_factor", C.c_double),
                ("ms_solve", C.c_double), ("factor_flops", C.c_double), ("solve_bytes", C.c_double),
                ("ops_per_round", C.c_double), ("body_ops", C.c_long), ("pairs", C.c_int),
                ("pairs_max", C.c_int), ("tune_s", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrainStats(C.Structure):
    _fields_ = [("steps", C.c_long), ("loss_first", C.c_double), ("loss_last", C.c_double),
                ("loss_recent", C.c_double)]


ABI = {
    "cb_last_error": (C.c_char_p, []),
    "cb_open": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "cb_close": (None, [C.c_void_p]),
    "cb_get_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "cb_get_dofmaps": (C.c_int, [C.c_void_p, PL, PL, PL]),
    "cb_get_joint_numbers": (C.c_int, [C.c_void_p, PL]),
    "cb_model_info": (C.c_int, [C.c_char_p, C.c_void_p, PL, PD, PL]),
    "cb_get_pattern": (C.c_int, [C.c_void_p, PI32, PI32]),
    "cb_get_load": (C.c_int, [C.c_void_p, PD]),
    "cb_state_set": (C.c_int, [C.c_void_p, C.c_char_p, PD, C.c_long]),
    "cb_state_get": (C.c_int, [C.c_void_p, C.c_char_p, PD, C.c_long]),
    "cb_tangent": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "cb_get_tangent": (C.c_int, [C.c_void_p, PD]),
    "cb_tangent_fp32": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "cb_internal_force": (C.c_int, [C.c_void_p, C.c_int, PD, PD]),
    "cb_get_yield_flags": (C.c_int, [C.c_void_p, PI, PD]),
    "cb_set_yield_flags": (C.c_int, [C.c_void_p, PI]),
    "cb_update_geometry": (C.c_int, [C.c_void_p, C.c_int, PD]),
    "cb_factor": (C.c_int, [C.c_void_p, C.POINTER(Options), PI, PD]),
    "cb_solve": (C.c_int, [C.c_void_p, PD, PD]),
    "cb_run": (C.c_int, [C.c_void_p, C.POINTER(Options), C.POINTER(Stats)]),
    "cb_get_step_events": (C.c_int, [C.c_void_p, PI, PD]),
    "cb_get_reactions": (C.c_int, [C.c_void_p, C.c_int, PD]),
    "cb_get_step_residuals": (C.c_int, [C.c_void_p, PD, PD]),
    "cb_get_path_forces": (C.c_int, [C.c_void_p, PD]),
    "cb_write_results": (C.c_int, [C.c_void_p, C.c_char_p]),
    "cb_dots": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_long, PD]),
    "cb_dots_noalloc": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_long,
                                  C.c_void_p, C.c_void_p]),
    "cb_get_path": (C.c_int, [C.c_void_p, PD, PI, PD]),
    "cb_get_step_stats": (C.c_int, [C.c_void_p, PD, PD, PL]),
    "cb_get_kernel_times": (C.c_int, [C.c_void_p, C.c_int, C.c_int, PD, PL]),
    "cb_mass": (C.c_int, [C.c_void_p, PD]),
    "cb_learned_create": (C.c_int, [C.c_uint, C.POINTER(C.c_void_p)]),
    "cb_learned_free": (None, [C.c_void_p]),
    "cb_learned_nparams": (C.c_long, [C.c_void_p]),
    "cb_learned_get_params": (C.c_int, [C.c_void_p, PF]),
    "cb_learned_set_params": (C.c_int, [C.c_void_p, PF]),
    "cb_learned_attach": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cb_learned_set_objective": (C.c_int, [C.c_void_p, C.c_int]),
    "cb_learned_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_void_p]),
    "cb_learned_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, PD, PD, C.c_int, PD, PF]),
    "cb_learned_blocks": (C.c_int, [C.c_void_p, C.c_void_p, PF, PF, PL, PL]),
    "cb_write_results2": (C.c_int, [C.c_void_p, C.c_char_p]),
    "cb_batch_open": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_void_p)]),
    "cb_batch_close": (None, [C.c_void_p]),
    "cb_batch_count": (C.c_int, [C.c_void_p]),
    "cb_batch_get_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Info)]),
    "cb_batch_run": (C.c_int, [C.c_void_p, C.POINTER(Options), C.POINTER(BatchStats)]),
    "cb_batch_get_status": (C.c_int, [C.c_void_p, PI, PI]),
    "cb_batch_get_tuning": (C.c_int, [C.c_void_p, PI, PD, C.c_int]),
    "cb_batch_get_path": (C.c_int, [C.c_void_p, C.c_int, PD, PI, PD]),
}


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python cu-bens_amd/build.py")
        _lib = C.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a: np.ndarray, kind=PD):
    return a.ctypes.data_as(kind)


class CubensError(RuntimeError):
    pass


def _check(rc: int):
    if rc != 0:
        raise CubensError(lib().cb_last_error().decode(errors="replace") or f"error {rc}")


def dots(a_ptrs, b_ptrs, length: int) -> np.ndarray:
    """cb_dots over device vectors given by their addresses (hipMalloc'd
    through the HIP runtime this library links); returns the n dot products."""
    n = len(a_ptrs)
    A = (C.c_void_p * n)(*a_ptrs)
    B = (C.c_void_p * n)(*b_ptrs)
    out = np.zeros(n)
    _check(lib().cb_dots(n, A, B, length, _p(out)))
    return out


def model_info(path: str) -> dict:
    """A model_def.txt read on the host alone (cb_model_info): sizes, the DOF
    map jcode [NJ*7], coordinates x [NJ*3] and the renumbering jnt [NJ]."""
    inf = Info()
    _check(lib().cb_model_info(path.encode(), C.byref(inf), None, None, None))
    jc = np.zeros(inf.nj * 7, np.int64)
    x = np.zeros(inf.nj * 3)
    jnt = np.zeros(inf.nj, np.int64)
    _check(lib().cb_model_info(path.encode(), C.byref(inf), _p(jc, PL), _p(x), _p(jnt, PL)))
    return {"info": inf, "nj": inf.nj, "neq": inf.neq, "algflag": inf.algflag, "jcode": jc, "x": x,
            "jnt": jnt}


class GpuModel:
    """A model opened on the GPU (cb_open)."""

    def __init__(self, path: str):
        h = C.c_void_p()
        _check(lib().cb_open(path.encode(), C.byref(h)))
        self.h = h
        inf = Info()
        _check(lib().cb_get_info(h, C.byref(inf)))
        self.info = inf
        self.neq = inf.neq
        self.nj = inf.nj
        self.ntr, self.nfr, self.nsh = inf.ne_tr, inf.ne_fr, inf.ne_sh

    def close(self):
        if getattr(self, "h", None):
            lib().cb_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- maps --------------------------------------------------------------
    def dofmaps(self):
        jc = np.zeros(self.nj * 7, np.int64)
        mc = np.zeros(max(1, self.ntr * 6 + self.nfr * 14 + self.nsh * 18), np.int64)
        mx = np.zeros(self.neq + 1, np.int64)
        _check(lib().cb_get_dofmaps(self.h, _p(jc, PL), _p(mc, PL), _p(mx, PL)))
        return jc, mc[: self.ntr * 6 + self.nfr * 14 + self.nsh * 18], mx

    def joint_numbers(self) -> np.ndarray:
        """jnt: input joint i is joint jnt[i] of the model (identity unless OPTFLAG 2)."""
        jnt = np.zeros(self.nj, np.int64)
        _check(lib().cb_get_joint_numbers(self.h, _p(jnt, PL)))
        return jnt

    def pattern(self):
        rp = np.zeros(self.nj + 1, np.int32)
        ci = np.zeros(self.info.nnz_blocks, np.int32)
        _check(lib().cb_get_pattern(self.h, _p(rp, PI32), _p(ci, PI32)))
        return rp, ci

    def load(self):
        q = np.zeros(self.neq)
        _check(lib().cb_get_load(self.h, _p(q)))
        return q

    # ---- state -------------------------------------------------------------
    def set(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        _check(lib().cb_state_set(self.h, name.encode(), _p(a), a.size))

    def get(self, name: str, n: int) -> np.ndarray:
        a = np.zeros(n)
        _check(lib().cb_state_get(self.h, name.encode(), _p(a), n))
        return a

    # ---- operations --------------------------------------------------------
    def tangent(self, gen_x: int = GEN_BASE, gen_state: int = GEN_BASE) -> np.ndarray:
        _check(lib().cb_tangent(self.h, gen_x, gen_state))
        v = np.zeros(self.info.nnz_blocks * self.info.bs * self.info.bs)
        _check(lib().cb_get_tangent(self.h, _p(v)))
        return v

    def tangent_fp32(self, gen_x: int = GEN_BASE, gen_state: int = GEN_BASE) -> np.ndarray:
        """The FP32 precision mode's tangent (cb_tangent_fp32), float32 joint blocks."""
        v = np.zeros(self.info.nnz_blocks * self.info.bs * self.info.bs, np.float32)
        _check(lib().cb_tangent_fp32(self.h, gen_x, gen_state, v.ctypes.data_as(C.POINTER(C.c_float))))
        return v

    def internal_force(self, dd: np.ndarray, gen_x: int = GEN_I, with_code: bool = False):
        """f of forces_tr/fr/sh; with_code: (f, forces_fr's code) under ANAFLAG 3."""
        dd = np.ascontiguousarray(dd, np.float64)
        f = np.zeros(self.neq)
        rc = lib().cb_internal_force(self.h, gen_x, _p(dd), _p(f))
        if rc < 0:
            _check(rc)
        if with_code:
            return f, rc
        _check(rc)
        return f

    def yield_flags(self):
        """(yldflag [NE_FR*2], regula-falsi fraction of the last code-1 force pass)."""
        y = np.zeros(max(self.nfr * 2, 1), np.int32)
        tau = np.zeros(1)
        _check(lib().cb_get_yield_flags(self.h, _p(y, PI), _p(tau)))
        return y[: self.nfr * 2], float(tau[0])

    def set_yield_flags(self, y):
        y = np.ascontiguousarray(y, np.int32)
        _check(lib().cb_set_yield_flags(self.h, _p(y, PI)))

    def mass(self) -> np.ndarray:
        """Lumped diagonal masses at the converged coordinates (cb_mass)."""
        sm = np.zeros(self.neq)
        _check(lib().cb_mass(self.h, _p(sm)))
        return sm

    def update_geometry(self, dd: np.ndarray, gen_x: int = GEN_I):
        dd = np.ascontiguousarray(dd, np.float64)
        _check(lib().cb_update_geometry(self.h, gen_x, _p(dd)))

    def factor(self, opt: Options | None = None):
        det = C.c_int()
        piv = np.zeros(self.neq)
        _check(lib().cb_factor(self.h, C.byref(opt) if opt is not None else None, C.byref(det), _p(piv)))
        return det.value, piv

    def solve(self, rhs: np.ndarray) -> np.ndarray:
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros(self.neq)
        _check(lib().cb_solve(self.h, _p(rhs), _p(x)))
        return x

    def run(self, opt: Options | None = None) -> dict:
        """Runs the driver; a driver that stops with an error (cb_run's positive
        return) gives its status and message, with the path up to there."""
        st = Stats()
        rc = lib().cb_run(self.h, C.byref(opt) if opt is not None else None, C.byref(st))
        msg = lib().cb_last_error().decode(errors="replace") if rc else ""
        if rc < 0:
            _check(rc)
        if rc != st.status:
            raise CubensError(f"cb_run returned {rc} with driver status {st.status}")
        n = st.nsteps
        lpf = np.zeros(max(n, 1))
        ite = np.zeros(max(n, 1), np.int32)
        keep = opt is None or opt.keep_d
        d = np.zeros((max(n, 1), self.neq)) if keep else None
        _check(lib().cb_get_path(self.h, _p(lpf), _p(ite, PI), _p(d) if keep else None))
        tw = np.zeros(max(n, 1))
        tj = np.zeros(max(n, 1))
        nw = np.zeros(max(n, 1), np.int64)
        _check(lib().cb_get_step_stats(self.h, _p(tw), _p(tj), _p(nw, PL)))
        ev = np.zeros((max(n, 1), NEV), np.int32)
        psi = np.zeros(max(n, 1))
        _check(lib().cb_get_step_events(self.h, _p(ev, PI), _p(psi)))
        res = np.zeros(max(n, 1))
        load = np.zeros(max(n, 1))
        _check(lib().cb_get_step_residuals(self.h, _p(res), _p(load)))
        out = st.as_dict()
        out.update(step_t=tw[:n], step_judge=tj[:n], step_newton=nw[:n], events=ev[:n], psi=psi[:n],
                   res=res[:n], load=load[:n])
        out.update(lpf=lpf[:n], itecnt=ite[:n], d=d[:n] if keep else None, msg=msg)
        if opt is not None and opt.keep_forces:
            nf = self.ntr + 7 * self.nfr + 6 * self.nsh
            fa = np.zeros(max(n * nf, 1))
            _check(lib().cb_get_path_forces(self.h, _p(fa)))
            out["forces"] = fa[: n * nf].reshape(n, nf)
        return out

    def reactions(self, gen: int = GEN_BASE) -> np.ndarray:
        """Support reactions [NJ, 7] of element-state generation gen (cb_get_reactions)."""
        r = np.zeros(self.nj * 7)
        _check(lib().cb_get_reactions(self.h, gen, _p(r)))
        return r.reshape(self.nj, 7)

    def kernel_times(self, precision: int = 64, step: int = -1) -> dict:
        nf = len(KT_NAMES)
        ms = np.zeros(nf)
        n = np.zeros(nf, np.int64)
        _check(lib().cb_get_kernel_times(self.h, precision, step, _p(ms), _p(n, PL)))
        return {KT_NAMES[f]: (float(ms[f]), int(n[f])) for f in range(nf)}

    def write_results2(self, path: str):
        _check(lib().cb_write_results2(self.h, path.encode()))

    def write_results(self, directory: str):
        """results2..5.txt of the last run in `directory` (cb_write_results)."""
        _check(lib().cb_write_results(self.h, directory.encode()))


class GpuBatch:
    """Many frame models of the nonlinear Newmark driver run side by side
    (cb_batch_open); each model's path is the one cb_run gives for it alone."""

    def __init__(self, paths):
        arr = (C.c_char_p * len(paths))(*[p.encode() for p in paths])
        h = C.c_void_p()
        _check(lib().cb_batch_open(arr, len(paths), C.byref(h)))
        self.h = h
        self.count = lib().cb_batch_count(h)
        self.info = []
        for i in range(self.count):
            inf = Info()
            _check(lib().cb_batch_get_info(h, i, C.byref(inf)))
            self.info.append(inf)

    def close(self):
        if getattr(self, "h", None):
            lib().cb_batch_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_steps: int = 1 << 30, keep_d: bool = True, timing: bool = False,
            graph: bool = False, autotune: bool = False, fp32_elements: bool = False) -> dict:
        """cb_batch_run; graph replays each round's Newton iteration as one HIP
        graph; autotune tunes the band factor's helper pairs on the first run;
        fp32_elements forms the member tangents in FP32."""
        st = BatchStats()
        opt = options(max_steps=max_steps, keep_d=int(keep_d), kernel_timing=int(timing), graph=int(graph),
                      autotune=int(autotune), fp32_elements=int(fp32_elements))
        _check(lib().cb_batch_run(self.h, C.byref(opt), C.byref(st)))
        status = np.zeros(self.count, np.int32)
        nsteps = np.zeros(self.count, np.int32)
        _check(lib().cb_batch_get_status(self.h, _p(status, PI), _p(nsteps, PI)))
        out = st.as_dict()
        out.update(status=status, nsteps=nsteps)
        return out

    def tuning(self) -> list:
        """The tuner's trials: [(helper pairs, factor ms)] (cb_batch_get_tuning)."""
        pairs = np.zeros(16, np.int32)
        ms = np.zeros(16)
        n = lib().cb_batch_get_tuning(self.h, _p(pairs, PI), _p(ms), 16)
        if n < 0:
            _check(n)
        return [(int(pairs[i]), float(ms[i])) for i in range(min(n, 16))]

    def path(self, i: int, with_d: bool = True) -> dict:
        """Model i's recorded steps: time, itecnt and (with_d) d [nsteps, NEQ_i]."""
        nsteps = np.zeros(self.count, np.int32)
        _check(lib().cb_batch_get_status(self.h, None, _p(nsteps, PI)))
        n = int(nsteps[i])
        t = np.zeros(max(n, 1))
        ite = np.zeros(max(n, 1), np.int32)
        d = np.zeros((max(n, 1), self.info[i].neq)) if with_d else None
        _check(lib().cb_batch_get_path(self.h, i, _p(t), _p(ite, PI), _p(d) if with_d else None))
        return {"lpf": t[:n], "itecnt": ite[:n], "d": d[:n] if with_d else None}


def bsr_to_dense(vals: np.ndarray, rowptr: np.ndarray, colidx: np.ndarray, bs: int,
                 eq_of_slot: np.ndarray, neq: int) -> np.ndarray:
    """Dense NEQ x NEQ matrix from joint-block values (fixed slots dropped)."""
    K = np.zeros((neq, neq))
    nj = len(rowptr) - 1
    v = vals.reshape(-1, bs, bs)
    for a in range(nj):
        for k in range(rowptr[a], rowptr[a + 1]):
            b = colidx[k]
            ia = eq_of_slot[a * bs:(a + 1) * bs]
            ib = eq_of_slot[b * bs:(b + 1) * bs]
            for p in range(bs):
                if ia[p] < 0:
                    continue
                for q in range(bs):
                    if ib[q] < 0:
                        continue
                    K[ia[p], ib[q]] = v[k, p, q]
    return K


def eq_of_slot(jcode: np.ndarray, nj: int, bs: int) -> np.ndarray:
    j = jcode.reshape(nj, 7)[:, :bs].astype(np.int64) - 1
    return j.reshape(-1)


class LearnedPrecond:
    """The learned block approximate-inverse preconditioner (cb_learned_*)."""

    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        _check(lib().cb_learned_create(seed, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().cb_learned_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def nparams(self) -> int:
        return int(lib().cb_learned_nparams(self.h))

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        _check(lib().cb_learned_get_params(self.h, _p(p, PF)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError("parameter vector size mismatch")
        _check(lib().cb_learned_set_params(self.h, _p(p, PF)))

    def set_objective(self, kind: int):
        _check(lib().cb_learned_set_objective(self.h, kind))

    def attach(self, g: "GpuModel"):
        _check(lib().cb_learned_attach(g.h, self.h))

    def train(self, g: "GpuModel", max_steps: int, probes: int = 4, lr: float = 1e-3) -> dict:
        st = TrainStats()
        _check(lib().cb_learned_train(self.h, g.h, max_steps, probes, lr, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def loss_grad(self, g: "GpuModel", r: np.ndarray, u: np.ndarray):
        r = np.ascontiguousarray(np.atleast_2d(r), np.float64)
        u = np.ascontiguousarray(np.atleast_2d(u), np.float64)
        loss = C.c_double()
        grad = np.zeros(self.nparams, np.float32)
        _check(lib().cb_learned_loss_grad(self.h, g.h, _p(r), _p(u), r.shape[0], C.byref(loss), _p(grad, PF)))
        return loss.value, grad

    def blocks(self, g: "GpuModel"):
        nv, ne = C.c_long(), C.c_long()
        _check(lib().cb_learned_blocks(self.h, g.h, None, None, C.byref(nv), C.byref(ne)))
        gd = np.zeros(nv.value * 36, np.float32)
        go = np.zeros(max(ne.value, 1) * 36, np.float32)
        _check(lib().cb_learned_blocks(self.h, g.h, _p(gd, PF), _p(go, PF), C.byref(nv), C.byref(ne)))
        return gd.reshape(-1, 6, 6), go[: ne.value * 36].reshape(-1, 6, 6)
