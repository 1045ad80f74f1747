"""ctypes view of the oracle (liboracle.so) and of the reference routines built
from source (_ref/libcubens_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libcubens_ref.so")

PD = C.POINTER(C.c_double)
PL = C.POINTER(C.c_long)
PI = C.POINTER(C.c_int)


class Dims(C.Structure):
    _fields_ = [("NJ", C.c_long), ("NE_TR", C.c_long), ("NE_FR", C.c_long),
                ("NE_SH", C.c_long), ("NEQ", C.c_long), ("ANAFLAG", C.c_int),
                ("ALGFLAG", C.c_int), ("SLVFLAG", C.c_int), ("OPTFLAG", C.c_int)]


class Model(C.Structure):
    _fields_ = [("dm", Dims), ("chkpt", C.c_int), ("rflag", C.c_int),
                ("minc", PL), ("jcode", PL), ("mcode", PL), ("maxa", PL), ("kht", PL),
                ("jnt", PL), ("wrpres", PI), ("lss", C.c_long), ("x", PD),
                ("emod", PD), ("yield_", PD), ("carea", PD), ("llength", PD), ("dens_tr", PD),
                ("gmod", PD), ("dens_fr", PD), ("istrong", PD), ("iweak", PD), ("ipolar", PD),
                ("iwarp", PD), ("zstrong", PD), ("zweak", PD),
                ("offset", PD), ("auxpt", PD), ("xfr", PD), ("osflag", PI), ("mendrel", PI),
                ("nu", PD), ("thick", PD), ("dens_sh", PD), ("xlocal", PD), ("farea", PD),
                ("slength", PD), ("dens", PD), ("c1", PD), ("c2", PD), ("c3", PD), ("q", PD),
                ("ntstpsinpt", C.c_long), ("ttot", C.c_double), ("dt", C.c_double),
                ("pinpt", PD), ("um", PD), ("vm", PD), ("am", PD),
                ("lpfmax", C.c_double), ("lpf0", C.c_double), ("dlpf0", C.c_double),
                ("dlpfmax", C.c_double), ("dlpfmin", C.c_double),
                ("itemax", C.c_int), ("submax", C.c_int), ("solmin", C.c_int),
                ("imagmax", C.c_int), ("negmax", C.c_int), ("iteopt", C.c_int),
                ("toldisp", C.c_double), ("tolforc", C.c_double), ("tolener", C.c_double),
                ("alpha", C.c_double), ("psi_thresh", C.c_double), ("dkimax", C.c_double),
                ("dk", C.c_double), ("dkdof", C.c_long),
                ("numopt", C.c_int), ("spectrds", C.c_double)]


class Path(C.Structure):
    _fields_ = [("max_steps", C.c_int), ("nsteps", C.c_int), ("lpf", PD), ("itecnt", PI),
                ("d", PD), ("f", PD), ("status", C.c_int), ("msg", C.c_char * 160),
                ("t_tangent", C.c_double), ("t_factor", C.c_double), ("t_solve", C.c_double),
                ("t_forces", C.c_double), ("t_other", C.c_double),
                ("n_factor", C.c_long), ("n_solve", C.c_long), ("n_newton", C.c_long),
                ("events", PI), ("psi", PD), ("react", PD), ("resid", PD), ("ef", PD),
                ("ckpt_path", C.c_char_p), ("ckpt_exact", C.c_int)]


class Ckpt(C.Structure):
    """orc_ckpt: the state of a Newmark checkpoint file (misc.c:494-716)."""
    _fields_ = [("tstep", C.c_long), ("uc", PD), ("vc", PD), ("ac", PD), ("ss", PD), ("sm", PD),
                ("d", PD), ("f", PD), ("ef", PD), ("x", PD), ("c1", PD), ("c2", PD), ("c3", PD),
                ("defllen", PD), ("llength", PD), ("efFE", PD), ("xfr", PD), ("yldflag", PI),
                ("deffarea", PD), ("defslen", PD), ("has_ext", C.c_int), ("ddt", C.c_double),
                ("llen_w", PD), ("farea_w", PD), ("slen_w", PD)]

# Arrays of a checkpoint file in file order, with their lengths for a model
# (NEQ, lss, NJ, element counts); the extension arrays last.
CKPT_ARRAYS = ("uc", "vc", "ac", "ss", "sm", "d", "f", "ef", "x", "c1", "c2", "c3", "defllen",
               "llength", "efFE", "xfr", "yldflag", "deffarea", "defslen", "llen_w", "farea_w",
               "slen_w")

# Slots of the per-step driver events (ORC_EV_* in cb_oracle.h).
EV_DET, EV_IMAG, EV_NEG, EV_DNORM, EV_HALVE, EV_YIELD, EV_UNLOAD = range(7)
NEV = 7


class Sample(C.Structure):
    _fields_ = [("t_tangent", C.c_double), ("t_forces", C.c_double), ("t_solve", C.c_double),
                ("t_factor_part", C.c_double), ("factor_fraction", C.c_double),
                ("t_factor_est", C.c_double), ("cols_done", C.c_long)]


class Backend(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("stiff_tr", C.c_void_p), ("forces_tr", C.c_void_p),
                ("stiff_sh", C.c_void_p), ("forces_sh", C.c_void_p), ("updatc", C.c_void_p),
                ("stiff_fr", C.c_void_p), ("forces_fr", C.c_void_p), ("mass_tr", C.c_void_p),
                ("mass_fr", C.c_void_p), ("mass_sh", C.c_void_p)]


def ptr(a: np.ndarray, kind=PD):
    return a.ctypes.data_as(kind)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_open.restype = C.POINTER(Model)
        _lib.orc_open.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        _lib.orc_free.argtypes = [C.POINTER(Model)]
        _lib.orc_run.argtypes = [C.POINTER(Model), C.POINTER(Backend), C.POINTER(Path)]
        _lib.orc_run.restype = C.c_int
        _lib.orc_ckpt_read.argtypes = [C.c_char_p, C.POINTER(Dims), C.c_long, C.POINTER(Ckpt)]
        _lib.orc_ckpt_read.restype = C.c_int
        _lib.orc_ckpt_write.argtypes = [C.c_char_p, C.POINTER(Dims), C.c_long, C.POINTER(Ckpt), C.c_int]
        _lib.orc_ckpt_write.restype = C.c_int
        _lib.orc_skyfact.argtypes = [C.c_long, PL, PD, PD, C.c_int, PI]
        _lib.orc_skyfact.restype = C.c_int
        _lib.orc_skysolve.argtypes = [C.c_long, PL, PD, PD]
        _lib.orc_quad.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, PD, PD, PD, PD,
                                  PD, PD, PD]
        _lib.orc_quad.restype = C.c_int
        _lib.orc_test.argtypes = [C.c_long, PD, PD, PD, PD, PD, PD, C.c_double, PI, C.c_double,
                                  C.c_double, C.c_double]
        _lib.orc_test.restype = C.c_int
        _lib.orc_time_sample.argtypes = [C.POINTER(Model), C.c_double, C.POINTER(Sample)]
        _lib.orc_time_sample.restype = C.c_int
        _lib.orc_shell_elem_tangent.argtypes = [C.c_double, C.c_double, C.c_double, PD,
                                                C.c_double, PD, C.c_double, PD, PD, PD, PD,
                                                C.c_int, PD]
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_PATH}")
        _ref = C.CDLL(REF_PATH)
        _ref.ref_open.argtypes = [C.c_char_p, C.POINTER(Dims), PL]
        _ref.ref_open.restype = C.c_int
        _ref.ref_get_long.argtypes = [C.c_char_p, PL, C.c_long]
        _ref.ref_get_double.argtypes = [C.c_char_p, PD, C.c_long]
        _ref.ref_fill_backend.argtypes = [C.POINTER(Backend)]
        _ref.ref_quad.argtypes = [C.POINTER(Dims), C.c_double, C.c_double, C.c_double, PD, PD,
                                  PD, PD, PD, PD, PD]
        _ref.ref_quad.restype = C.c_int
        _ref.ref_output_files.argtypes = [C.c_char_p, C.c_int, PD, PI, PD, PD]
        _ref.ref_checkpoint.argtypes = [C.c_char_p, C.c_long, C.c_long] + [PD] * 16 + [PI, PD, PD]
        _ref.ref_checkpoint.restype = C.c_int
        _ref.ref_restart.argtypes = [C.c_char_p, PL, C.c_long] + [PD] * 16 + [PI, PD, PD]
        _ref.ref_restart.restype = C.c_int
        _ref.ref_output_files.restype = C.c_int
        _ref.ref_test.argtypes = [C.POINTER(Dims), PD, PD, PD, PD, PD, PD, C.c_double, PI,
                                  C.c_double, C.c_double, C.c_double]
        _ref.ref_test.restype = C.c_int
    return _ref


def _arr(p, n, dtype=np.float64):
    if n <= 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


class OracleModel:
    """A parsed model held by the oracle, with numpy copies of its arrays."""

    def __init__(self, path: str):
        err = C.create_string_buffer(256)
        self.path = path
        self.h = lib().orc_open(path.encode(), err, 256)
        if not self.h:
            raise ValueError(err.value.decode())
        m = self.h.contents
        d = m.dm
        self.dims = Dims(d.NJ, d.NE_TR, d.NE_FR, d.NE_SH, d.NEQ, d.ANAFLAG, d.ALGFLAG,
                         d.SLVFLAG, d.OPTFLAG)
        self.nj, self.ntr, self.nfr, self.nsh, self.neq = d.NJ, d.NE_TR, d.NE_FR, d.NE_SH, d.NEQ
        self.lss = m.lss
        ntri = self.ntr + 3 * self.nfr + 3 * self.nsh
        self.jcode = _arr(m.jcode, self.nj * 7, np.int64)
        self.mcode = _arr(m.mcode, self.ntr * 6 + self.nfr * 14 + self.nsh * 18, np.int64)
        self.minc = _arr(m.minc, self.ntr * 2 + self.nfr * 2 + self.nsh * 3, np.int64)
        self.maxa = _arr(m.maxa, self.neq + 1, np.int64)
        self.kht = _arr(m.kht, self.neq, np.int64)
        self.x = _arr(m.x, self.nj * 3)
        self.q = _arr(m.q, self.neq)
        self.emod = _arr(m.emod, self.ntr + self.nfr + self.nsh)
        self.yield_ = _arr(m.yield_, self.ntr + self.nfr + self.nsh)
        self.zstrong = _arr(m.zstrong, self.nfr)
        self.zweak = _arr(m.zweak, self.nfr)
        self.anaflag = m.dm.ANAFLAG
        self.c1 = _arr(m.c1, ntri)
        self.c2 = _arr(m.c2, ntri)
        self.c3 = _arr(m.c3, ntri)
        self.nu = _arr(m.nu, self.nsh)
        self.thick = _arr(m.thick, self.nsh)
        self.xlocal = _arr(m.xlocal, self.nsh * 3)
        self.farea = _arr(m.farea, self.nsh)
        self.slength = _arr(m.slength, self.nsh * 3)
        self.carea = _arr(m.carea, self.ntr + self.nfr)
        self.llength = _arr(m.llength, self.ntr + self.nfr)
        self.dkdof = m.dkdof
        self.auxpt = _arr(m.auxpt, self.nfr * 3)
        self.offset = _arr(m.offset, self.nfr * 6)
        self.osflag = _arr(m.osflag, self.nfr, np.int32)

    def close(self):
        if self.h:
            lib().orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def time_sample(self, factor_budget_s: float) -> dict:
        """Bounded single-core timing sample of the serial path (orc_time_sample)."""
        smp = Sample()
        if lib().orc_time_sample(self.h, factor_budget_s, C.byref(smp)):
            raise MemoryError("oracle sample allocation failed")
        return {k: getattr(smp, k) for k, _ in smp._fields_}

    def reactions(self, ef, c1, c2, c3) -> np.ndarray:
        """Support reactions [NJ, 7] of an element state (orc_reactions)."""
        L = lib()
        L.orc_reactions.argtypes = [C.POINTER(Model), PD, PD, PD, PD, PD]
        r = np.zeros(self.nj * 7)
        L.orc_reactions(self.h, _c(ef), _c(c1), _c(c2), _c(c3), ptr(r))
        return r.reshape(self.nj, 7)

    def run(self, max_steps: int, backend: str = "oracle", keep_f: bool = False,
            keep_react: bool = False, keep_ef: bool = False, ckpt_path: str | None = None,
            ckpt_exact: bool = False) -> dict:
        """Runs the analysis; returns the converged path and timings (and, with
        keep_react, the reactions [nsteps, NJ, 7] of every converged step).
        ckpt_path: the Newmark checkpoint file, written every CHKPT steps and
        read first when the model has RFLAG 1."""
        P = Path()
        P.max_steps = max_steps
        P.ckpt_path = ckpt_path.encode() if ckpt_path else None
        P.ckpt_exact = int(bool(ckpt_exact))
        lpf = np.zeros(max_steps)
        ite = np.zeros(max_steps, np.int32)
        dmat = np.zeros((max_steps, max(self.neq, 1)))
        fmat = np.zeros((max_steps, max(self.neq, 1))) if keep_f else None
        P.lpf = ptr(lpf)
        P.itecnt = ptr(ite, PI)
        P.d = ptr(dmat)
        P.f = ptr(fmat) if keep_f else None
        ev = np.zeros((max_steps, NEV), np.int32)
        psi = np.zeros(max_steps)
        P.events = ptr(ev, PI)
        P.psi = ptr(psi)
        react = np.zeros((max_steps, self.nj, 7)) if keep_react else None
        P.react = ptr(react) if keep_react else None
        resid = np.zeros((max_steps, 2))
        P.resid = ptr(resid)
        nef = self.ntr * 2 + self.nfr * 14 + self.nsh * 18
        efp = np.zeros((max_steps, max(nef, 1))) if keep_ef else None
        P.ef = ptr(efp) if keep_ef else None
        be = None
        if backend == "ref":
            be = Backend()
            ref().ref_fill_backend(C.byref(be))
        rc = lib().orc_run(self.h, C.byref(be) if be is not None else None, C.byref(P))
        n = P.nsteps
        out = {"rc": rc, "status": P.status, "msg": P.msg.decode(), "nsteps": n,
               "lpf": lpf[:n].copy(), "itecnt": ite[:n].copy(), "d": dmat[:n].copy(),
               "t_tangent": P.t_tangent, "t_factor": P.t_factor, "t_solve": P.t_solve,
               "t_forces": P.t_forces, "n_factor": P.n_factor, "n_solve": P.n_solve,
               "n_newton": P.n_newton, "events": ev[:n].copy(), "psi": psi[:n].copy(),
               "res": resid[:n, 0].copy(), "load": resid[:n, 1].copy()}
        if keep_f:
            out["f"] = fmat[:n].copy()
        if keep_react:
            out["react"] = react[:n].copy()
        if keep_ef:
            out["ef"] = efp[:n].copy()
        return out


def ckpt_sizes(om) -> dict:
    """Length of each checkpoint array of model om (orc_ckpt layouts)."""
    ntri = om.ntr + 3 * om.nfr + 3 * om.nsh
    return {"uc": om.neq, "vc": om.neq, "ac": om.neq, "ss": om.lss, "sm": om.lss, "d": om.neq,
            "f": om.neq, "ef": om.ntr * 2 + om.nfr * 14 + om.nsh * 18, "x": om.nj * 3, "c1": ntri,
            "c2": ntri, "c3": ntri, "defllen": om.ntr + om.nfr, "llength": om.ntr + om.nfr,
            "efFE": om.nfr * 14, "xfr": om.nfr * 6, "yldflag": om.nfr * 2, "deffarea": om.nsh,
            "defslen": om.nsh * 3, "llen_w": om.ntr + om.nfr, "farea_w": om.nsh,
            "slen_w": om.nsh * 3}


def _ckpt_struct(om):
    arrs = {k: np.zeros(max(n, 1), np.int32 if k == "yldflag" else np.float64)
            for k, n in ckpt_sizes(om).items()}
    c = Ckpt()
    for k, a in arrs.items():
        setattr(c, k, ptr(a, PI) if k == "yldflag" else ptr(a))
    return c, arrs


def read_checkpoint(path: str, om) -> dict:
    """The oracle's reading of a checkpoint file (orc_ckpt_read) for model om:
    tstep, every array at its length, and the extension (has_ext, ddt)."""
    c, arrs = _ckpt_struct(om)
    rc = lib().orc_ckpt_read(path.encode(), C.byref(om.h.contents.dm), om.lss, C.byref(c))
    if rc:
        raise ValueError(f"checkpoint file not readable ({rc})")
    out = {k: arrs[k][:n].copy() for k, n in ckpt_sizes(om).items()}
    out.update(tstep=c.tstep, has_ext=c.has_ext, ddt=c.ddt)
    return out


def write_checkpoint(path: str, om, state: dict, exact: bool = False):
    """orc_ckpt_write of a state dict shaped as read_checkpoint returns it."""
    c, arrs = _ckpt_struct(om)
    for k, n in ckpt_sizes(om).items():
        arrs[k][:n] = state[k][:n]
    c.tstep = int(state["tstep"])
    c.ddt = float(state.get("ddt", 1.0))
    if lib().orc_ckpt_write(path.encode(), C.byref(om.h.contents.dm), om.lss, C.byref(c), int(exact)):
        raise OSError(f"cannot write {path}")


class RefModel:
    """The reference's own input sequence run on a model_def.txt."""

    def __init__(self, path: str):
        d = Dims()
        lss = C.c_long()
        if ref().ref_open(path.encode(), C.byref(d), C.byref(lss)):
            raise ValueError("reference input sequence failed")
        self.dims = d
        self.nj, self.ntr, self.nfr, self.nsh, self.neq = d.NJ, d.NE_TR, d.NE_FR, d.NE_SH, d.NEQ
        self.lss = lss.value

    def get_long(self, name: str, n: int) -> np.ndarray:
        a = np.zeros(max(n, 1), np.int64)
        if ref().ref_get_long(name.encode(), ptr(a, PL), n):
            raise KeyError(name)
        return a[:n]

    def get_double(self, name: str, n: int) -> np.ndarray:
        a = np.zeros(max(n, 1))
        if ref().ref_get_double(name.encode(), ptr(a), n):
            raise KeyError(name)
        return a[:n]


def skyline_to_dense(ss: np.ndarray, maxa: np.ndarray) -> np.ndarray:
    """Dense symmetric matrix from a skyline array with 1-based maxa."""
    n = len(maxa) - 1
    K = np.zeros((n, n))
    for j in range(n):
        lo, hi = maxa[j], maxa[j + 1]
        for a in range(lo, hi):
            i = j - (a - lo)
            K[i, j] = K[j, i] = ss[a - 1]
    return K


def bind_elements():
    """ctypes signatures of the oracle's element routines (reference argument order)."""
    L = lib()
    D = C.POINTER(Dims)
    L.orc_stiff_sh.argtypes = [D] + [PD] * 12 + [PL] * 3
    L.orc_forces_sh.argtypes = [D, PD, PD, PD] + [PD] * 14 + [PL, PL]
    L.orc_updatc.argtypes = [D, PD, PD, PD, PD, PD, PD, PD, PD, PI, PD, PD, PD, PD, PL, PL]
    L.orc_stiff_tr.argtypes = [D, PD] + [PD] * 9 + [PL, PL]
    L.orc_forces_tr.argtypes = [D, PD, PD] + [PD] * 8 + [PL]
    L.orc_orbison.argtypes = [C.c_double] * 3
    L.orc_orbison.restype = C.c_double
    L.orc_orbison_grad.argtypes = [C.c_double] * 6 + [PD]
    L.orc_frame_plastic.argtypes = [PD, PD, PI, C.c_double, C.c_double, C.c_double]
    L.orc_regula_falsi.argtypes = [C.c_double] * 6
    L.orc_regula_falsi.restype = C.c_double
    L.orc_frame_unload.argtypes = [PD, PD, C.c_double, C.c_double, C.c_double, PD, PD]
    L.orc_frame_unload.restype = C.c_int
    return L


def orbison(p, my, mz) -> float:
    """Orbison's yield function of normalised end forces (frame.c:613-620)."""
    return bind_elements().orc_orbison(p, my, mz)


def orbison_grad(p, my, mz, Py, Mpy, Mpz) -> np.ndarray:
    g = np.zeros(3)
    bind_elements().orc_orbison_grad(p, my, mz, Py, Mpy, Mpz, ptr(g))
    return g


def frame_plastic(k, eftot, yld, Py, Mpy, Mpz) -> np.ndarray:
    """stiffm_fr (frame.c:580-797) on a copy of the 14x14 local tangent k."""
    out = np.array(k, np.float64, copy=True, order="C")
    y = np.ascontiguousarray(yld, np.int32)
    bind_elements().orc_frame_plastic(ptr(out), _c(eftot), ptr(y, PI), Py, Mpy, Mpz)
    return out


def regula_falsi(p, dp, my, dmy, mz, dmz) -> float:
    return bind_elements().orc_regula_falsi(p, dp, my, dmy, mz, dmz)


def frame_unload(phi, eftot_i, Py, Mpy, Mpz, k, dl) -> int:
    return bind_elements().orc_frame_unload(_c(phi), _c(eftot_i), Py, Mpy, Mpz, _c(k), _c(dl))


def stiff_sh(om: "OracleModel", x, c1, c2, c3, deffarea) -> np.ndarray:
    """Skyline array of stiff_sh at the given state (ss zeroed first)."""
    L = bind_elements()
    ss = np.zeros(om.lss)
    L.orc_stiff_sh(C.byref(om.dims), ptr(ss), ptr(om.emod), ptr(om.nu), ptr(np.ascontiguousarray(x)),
                   ptr(om.xlocal), ptr(om.thick), ptr(om.farea), ptr(np.ascontiguousarray(deffarea)),
                   ptr(om.slength), ptr(np.ascontiguousarray(c1)), ptr(np.ascontiguousarray(c2)),
                   ptr(np.ascontiguousarray(c3)), ptr(om.maxa, PL), ptr(om.minc, PL), ptr(om.mcode, PL))
    return ss


def stiff_tr(om: "OracleModel", c1, c2, c3, defllen, ef) -> np.ndarray:
    L = bind_elements()
    ss = np.zeros(om.lss)
    L.orc_stiff_tr(C.byref(om.dims), ptr(ss), ptr(om.emod), ptr(om.carea), ptr(om.llength),
                   ptr(np.ascontiguousarray(defllen)), ptr(om.yield_), ptr(np.ascontiguousarray(c1)),
                   ptr(np.ascontiguousarray(c2)), ptr(np.ascontiguousarray(c3)),
                   ptr(np.ascontiguousarray(ef)), ptr(om.maxa, PL), ptr(om.mcode, PL))
    return ss


def forces_sh(om: "OracleModel", dd, x_temp, c_ip, c_i, ef_ip):
    """forces_sh: returns (f, ef_i, ef_ip after the call)."""
    L = bind_elements()
    f = np.zeros(om.neq)
    efp = np.array(ef_ip, dtype=np.float64, copy=True)
    efi = np.zeros_like(efp)
    L.orc_forces_sh(C.byref(om.dims), ptr(f), ptr(efp), ptr(efi), ptr(np.ascontiguousarray(dd)),
                    ptr(np.ascontiguousarray(x_temp)), ptr(om.emod), ptr(om.nu), ptr(om.xlocal),
                    ptr(om.thick), ptr(om.farea), ptr(om.slength),
                    *[ptr(np.ascontiguousarray(v)) for v in (*c_ip, *c_i)],
                    ptr(om.minc, PL), ptr(om.mcode, PL))
    return f, efi, efp


def forces_tr(om: "OracleModel", c1, c2, c3, defllen):
    L = bind_elements()
    f = np.zeros(om.neq)
    ef = np.zeros(om.ntr * 2)
    L.orc_forces_tr(C.byref(om.dims), ptr(f), ptr(ef), ptr(om.emod), ptr(om.carea), ptr(om.llength),
                    ptr(np.ascontiguousarray(defllen)), ptr(om.yield_), ptr(np.ascontiguousarray(c1)),
                    ptr(np.ascontiguousarray(c2)), ptr(np.ascontiguousarray(c3)), ptr(om.mcode, PL))
    return f, ef


def updatc(om: "OracleModel", x_temp, dd):
    """updatc: returns (x_new, x_ip, defllen_i, deffarea_i, defslen_i, c1_i, c2_i, c3_i)."""
    L = bind_elements()
    x = np.array(x_temp, dtype=np.float64, copy=True)
    xip = np.zeros_like(x)
    nfr = om.nfr
    xfr = np.zeros(max(1, nfr * 6))
    defllen = np.zeros(max(1, om.ntr + nfr))
    deffarea = np.zeros(max(1, om.nsh))
    defslen = np.zeros(max(1, om.nsh * 3))
    ntri = om.ntr + 3 * nfr + 3 * om.nsh
    c1, c2, c3 = np.zeros(ntri), np.zeros(ntri), np.zeros(ntri)
    off = np.zeros(max(1, nfr * 6))
    osf = np.zeros(max(1, nfr), np.int32)
    aux = np.zeros(max(1, nfr * 3))
    if nfr:
        off[:nfr * 6] = om.offset
        osf[:nfr] = om.osflag
        aux[:nfr * 3] = om.auxpt
    L.orc_updatc(C.byref(om.dims), ptr(x), ptr(xip), ptr(xfr), ptr(np.ascontiguousarray(dd)),
                 ptr(defllen), ptr(deffarea), ptr(defslen), ptr(off), ptr(osf, PI), ptr(aux),
                 ptr(c1), ptr(c2), ptr(c3), ptr(om.minc, PL), ptr(om.jcode, PL))
    return x, xip, defllen[: om.ntr + nfr], deffarea[: om.nsh], defslen[: om.nsh * 3], c1, c2, c3


def skyfact(maxa: np.ndarray, ss: np.ndarray, msal: int = 1):
    """Returns (rc, factor, pivots, det)."""
    L = lib()
    a = np.array(ss, dtype=np.float64, copy=True)
    n = len(maxa) - 1
    piv = np.zeros(n)
    det = C.c_int()
    rc = L.orc_skyfact(n, ptr(np.ascontiguousarray(maxa, np.int64), PL), ptr(a), ptr(piv), msal, C.byref(det))
    return rc, a, piv, det.value


def skysolve(maxa: np.ndarray, fac: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    L = lib()
    x = np.array(rhs, dtype=np.float64, copy=True)
    L.orc_skysolve(len(maxa) - 1, ptr(np.ascontiguousarray(maxa, np.int64), PL), ptr(fac), ptr(x))
    return x


def frame_ke(E, G, A, L, Is, Iw, J, Cw) -> np.ndarray:
    """Local 14x14 elastic frame stiffness (orc_frame_ke)."""
    k = np.zeros((14, 14))
    L_ = lib()
    L_.orc_frame_ke.argtypes = [C.c_double] * 8 + [PD]
    L_.orc_frame_ke(E, G, A, L, Is, Iw, J, Cw, ptr(k))
    return k


def frame_kg(f, L, A, Ip) -> np.ndarray:
    """Local 14x14 geometric frame stiffness for end forces f[14] (orc_frame_kg)."""
    k = np.zeros((14, 14))
    f = np.ascontiguousarray(f, np.float64)
    L_ = lib()
    L_.orc_frame_kg.argtypes = [PD, C.c_double, C.c_double, C.c_double, PD]
    L_.orc_frame_kg(ptr(f), L, A, Ip, ptr(k))
    return k


def _c(a):
    return ptr(np.ascontiguousarray(a, np.float64))


def stiff_fr(om: "OracleModel", llength, defllen_ip, c1, c2, c3, ef_ip, yld=None) -> np.ndarray:
    """Skyline array of stiff_fr at the given state (ss zeroed first); yld:
    the yldflag state under ANAFLAG 3."""
    L_ = lib()
    L_.orc_stiff_fr.argtypes = [C.POINTER(Dims), PD, C.c_void_p] + [PD] * 7 + [PI]
    ss = np.zeros(om.lss)
    efFE = np.zeros(max(om.nfr * 14, 1))
    y = np.zeros(max(om.nfr * 2, 1), np.int32) if yld is None else np.ascontiguousarray(yld, np.int32)
    L_.orc_stiff_fr(C.byref(om.dims), ptr(ss), om.h, _c(llength), _c(defllen_ip), _c(c1), _c(c2),
                    _c(c3), _c(ef_ip), ptr(efFE), ptr(y, PI))
    return ss


def forces_fr(om: "OracleModel", dd, llength, defllen_ip, c_ip, c_i, ef_ip, yld=None, dlpf=1.0):
    """forces_fr: returns (f, ef_i), or under ANAFLAG 3 with yld given
    (f, ef_i, code, yld after the call, dlpf after the call)."""
    L_ = lib()
    L_.orc_forces_fr.argtypes = ([C.POINTER(Dims), PD, PD, PD, PD, PD, PD, C.c_void_p] + [PD] * 8 +
                                 [PI, PD])
    L_.orc_forces_fr.restype = C.c_int
    f = np.zeros(om.neq)
    efi = np.zeros(len(ef_ip))
    efFE = np.zeros(max(om.nfr * 14, 1))
    efFE_i = np.zeros_like(efFE)
    y = (np.zeros(max(om.nfr * 2, 1), np.int32) if yld is None
         else np.array(yld, np.int32, copy=True))
    dl = np.array([dlpf], np.float64)
    code = L_.orc_forces_fr(C.byref(om.dims), ptr(f), _c(ef_ip), ptr(efi), ptr(efFE), ptr(efFE_i),
                            _c(dd), om.h, _c(llength), _c(defllen_ip), *[_c(v) for v in (*c_ip, *c_i)],
                            ptr(y, PI), ptr(dl))
    if yld is None:
        return f, efi
    return f, efi, code, y, float(dl[0])


def mass(om: "OracleModel", x):
    """Lumped masses of every element type at coordinates x; returns
    (sm, llength, farea, slength) with the rewritten working geometry."""
    L_ = lib()
    D = C.POINTER(Dims)
    L_.orc_mass_tr.argtypes = [D, PD, C.c_void_p, PD, PD, PD]
    L_.orc_mass_fr.argtypes = [D, PD, C.c_void_p, PD, PD, PD, PD]
    L_.orc_mass_sh.argtypes = [D, PD, C.c_void_p, PD, PD, PD, PD]
    m = C.cast(om.h, C.POINTER(Model)).contents
    dens = _arr(m.dens, om.ntr + om.nfr + om.nsh)
    sm = np.zeros(om.neq)
    llen = om.llength.copy() if len(om.llength) else np.zeros(1)
    farea = om.farea.copy() if om.nsh else np.zeros(1)
    slen = om.slength.copy() if om.nsh else np.zeros(1)
    xfr = np.zeros(max(om.nfr * 6, 1))
    x = np.ascontiguousarray(x, np.float64)
    if om.ntr:
        L_.orc_mass_tr(C.byref(om.dims), ptr(sm), om.h, ptr(x), ptr(llen), ptr(dens))
    if om.nfr:
        L_.orc_mass_fr(C.byref(om.dims), ptr(sm), om.h, ptr(x), ptr(llen), ptr(xfr), ptr(dens))
    if om.nsh:
        L_.orc_mass_sh(C.byref(om.dims), ptr(sm), om.h, ptr(x), ptr(farea), ptr(slen), ptr(dens))
    return sm, llen, farea, slen
