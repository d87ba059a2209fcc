"""ctypes front end of the checker libraries (test infrastructure only).

* ``liboracle.so``         -- the C restatement in tjsq_oracle.c;
* ``_ref/libtjsq_ref.so``  -- the reference's own Fortran routines behind
  ref_bridge.f90 (present where oracle/Makefile could build it).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lock = threading.RLock()
_libs = {}

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Lattice(C.Structure):
    _fields_ = [("ny", C.c_int), ("nx_full", C.c_int), ("jt", C.c_double * 2), ("jd", C.c_double * 2),
                ("jn", C.c_double * 2), ("esite", C.c_double), ("up_bias", C.c_int)]


def lattice_struct(model) -> Lattice:
    lat = Lattice()
    lat.ny = model.ny
    lat.nx_full = model.table_columns
    lat.jt[0], lat.jt[1] = model.jt
    lat.jd[0], lat.jd[1] = model.jd
    lat.jn[0], lat.jn[1] = model.jn
    lat.esite = model.esite
    lat.up_bias = model.up_bias
    return lat


def _model_args():
    return [C.c_int, I32P, I32P, I64P, F64P]


def oracle_lib():
    with _lock:
        if "oracle" not in _libs:
            lib = C.CDLL(os.path.join(HERE, "liboracle.so"))
            lib.oracle_w6js.restype = C.c_double
            lib.oracle_w6js.argtypes = [C.c_int] * 6
            lib.oracle_model_wave.restype = C.c_int
            lib.oracle_model_wave.argtypes = ([C.POINTER(Lattice), C.c_int, C.c_int] + _model_args() + _model_args()
                                              + [C.c_int, I32P, I32P, C.c_int, I32P, I32P, C.c_int, I32P, F64P, F64P])
            lib.oracle_get_basis.restype = C.c_int
            lib.oracle_get_basis.argtypes = [C.c_int, C.c_int, I32P, C.c_int, I32P, I32P, I32P, C.POINTER(C.c_int)]
            lib.oracle_get_super_basis.restype = C.c_int
            lib.oracle_get_super_basis.argtypes = [C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_int, I32P]
            lib.oracle_hmerge_tables.restype = C.c_int
            lib.oracle_hmerge_tables.argtypes = [C.POINTER(Lattice), C.c_int, C.c_int, I32P, I32P, I32P]
            _libs["oracle"] = lib
        return _libs["oracle"]


def ref_path() -> str:
    return os.path.join(HERE, "_ref", "libtjsq_ref.so")


def ref_available() -> bool:
    return os.path.exists(ref_path())


def ref_lib():
    with _lock:
        if "ref" not in _libs:
            lib = C.CDLL(ref_path(), mode=C.RTLD_GLOBAL)
            lib.refb_init.restype = C.c_int
            lib.refb_init.argtypes = [C.c_int, C.c_int] + [C.c_double] * 8
            margs = [C.c_int, I32P, I32P, I64P, F64P]
            common = [C.c_int, C.c_int] + margs + margs + [C.c_int, I32P, I32P, C.c_int, I32P, I32P, C.c_int, I32P]
            lib.refb_model_wave.restype = C.c_int
            lib.refb_model_wave.argtypes = common + [F64P, F64P]
            lib.refb_lanczos.restype = C.c_int
            lib.refb_lanczos.argtypes = common + [F64P, C.POINTER(C.c_double)]
            lib.refb_w6js.restype = C.c_double
            lib.refb_w6js.argtypes = [C.c_int] * 6
            lib.refb_get_basis.restype = C.c_int
            lib.refb_get_basis.argtypes = [C.c_int, C.c_int, I32P, C.c_int, I32P, I32P, I32P, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
            _libs["ref"] = lib
        return _libs["ref"]


def _m(ops):
    return [ops.nops, np.ascontiguousarray(ops.hdr), np.ascontiguousarray(ops.sub), np.ascontiguousarray(ops.off),
            np.ascontiguousarray(ops.data)]


def _bs(bs):
    return [bs.nsec, np.ascontiguousarray(bs.sec), np.ascontiguousarray(bs.sub)]


def w6js(*j) -> float:
    return oracle_lib().oracle_w6js(*[int(v) for v in j])


def model_wave(model, sb, x: np.ndarray) -> np.ndarray:
    """Oracle y = H_eff x."""
    lib = oracle_lib()
    lat = lattice_struct(model)
    y = np.zeros(sb.dim)
    rc = lib.oracle_model_wave(C.byref(lat), sb.sys_len, sb.env_len, *_m(sb.sys), *_m(sb.env), *_bs(sb.sys_bs),
                               *_bs(sb.env_bs), int(sb.wsec.shape[0]), np.ascontiguousarray(sb.wsec),
                               np.ascontiguousarray(x, dtype=np.float64), y)
    if rc != 0:
        raise RuntimeError(f"oracle_model_wave failed: {rc}")
    return y


def hmerge_tables(model, sys_len, env_len):
    lib = oracle_lib()
    lat = lattice_struct(model)
    cbs = np.zeros(4, np.int32)
    cbb = np.zeros(2, np.int32)
    inb = np.zeros(sys_len + env_len + 3, np.int32)
    if lib.oracle_hmerge_tables(C.byref(lat), sys_len, env_len, cbs, cbb, inb) != 0:
        raise RuntimeError("oracle_hmerge_tables failed")
    return cbs, cbb, inb


def _ref_init(model, lanzero=1e-10):
    lib = ref_lib()
    rc = lib.refb_init(model.ny, model.up_bias, model.jt[0], model.jt[1], model.jd[0], model.jd[1], model.jn[0],
                       model.jn[1], model.esite, lanzero)
    if rc != 0:
        raise RuntimeError("reference build is compiled for a different lattice width")
    return lib


def ref_model_wave(model, sb, x: np.ndarray) -> np.ndarray:
    """y = H_eff x through the reference's own hmerge + model_wave."""
    with _lock:
        lib = _ref_init(model)
        y = np.zeros(sb.dim)
        lib.refb_model_wave(sb.sys_len, sb.env_len, *_m(sb.sys), *_m(sb.env), *_bs(sb.sys_bs), *_bs(sb.env_bs),
                            int(sb.wsec.shape[0]), np.ascontiguousarray(sb.wsec),
                            np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def ref_lanczos(model, sb, x0: np.ndarray, lanzero: float = 1e-10):
    """(E0, psi) through the reference's lanczos_memo_new with start vector x0.
    The reference appends to Energy_All.dat in the working directory."""
    with _lock:
        lib = _ref_init(model, lanzero)
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        e0 = C.c_double(0.0)
        lib.refb_lanczos(sb.sys_len, sb.env_len, *_m(sb.sys), *_m(sb.env), *_bs(sb.sys_bs), *_bs(sb.env_bs),
                         int(sb.wsec.shape[0]), np.ascontiguousarray(sb.wsec), x, C.byref(e0))
    return e0.value, x


_ref_ready = False


def ref_w6js(*j) -> float:
    """W6JS of the reference build."""
    global _ref_ready
    with _lock:
        lib = ref_lib()
        if not _ref_ready:
            lib.refb_init(6, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1e-10)
            _ref_ready = True
        return lib.refb_w6js(*[int(v) for v in j])


def ref_get_basis(block_len, blk, site):
    """Get_basis of the reference build: (sec, sub) tables."""
    blk = np.ascontiguousarray(blk, dtype=np.int32)
    site = np.ascontiguousarray(site, dtype=np.int32)
    cap = 3 * max(1, blk.shape[0]) * max(1, site.shape[0])
    sec = np.zeros(4 * cap, np.int32)
    sub = np.zeros(6 * cap, np.int32)
    ns, nq = C.c_int(0), C.c_int(0)
    with _lock:
        ref_lib().refb_get_basis(block_len, blk.shape[0], blk, site.shape[0], site, sec, sub, C.byref(ns),
                                 C.byref(nq))
    return sec[:4 * ns.value].reshape(-1, 4), sub[:6 * nq.value].reshape(-1, 6)


def oracle_get_basis(block_len, blk, site):
    blk = np.ascontiguousarray(blk, dtype=np.int32)
    site = np.ascontiguousarray(site, dtype=np.int32)
    cap = 3 * max(1, blk.shape[0]) * max(1, site.shape[0])
    sec = np.zeros(4 * cap, np.int32)
    sub = np.zeros(6 * cap, np.int32)
    nq = C.c_int(0)
    ns = oracle_lib().oracle_get_basis(block_len, blk.shape[0], blk, site.shape[0], site, sec, sub, C.byref(nq))
    return sec[:4 * ns].reshape(-1, 4), sub[:6 * nq.value].reshape(-1, 6)


def dropin_path() -> str:
    return os.path.join(HERE, "_ref", "libtjsq_dropin.so")


def dropin_lib():
    """The product's Fortran drop-in shim (grand_sql_tj_amd/fortran) built
    inside the reference build, with its test entry points."""
    with _lock:
        if "dropin" not in _libs:
            ref_lib()
            lib = C.CDLL(dropin_path(), mode=C.RTLD_GLOBAL)
            margs = [C.c_int, I32P, I32P, I64P, F64P]
            common = [C.c_int, C.c_int] + margs + margs + [C.c_int, I32P, I32P, C.c_int, I32P, I32P, C.c_int, I32P]
            lib.dropin_model_wave.restype = C.c_int
            lib.dropin_model_wave.argtypes = common + [F64P, F64P]
            lib.dropin_lanczos.restype = C.c_int
            lib.dropin_lanczos.argtypes = common + [F64P, C.POINTER(C.c_double)]
            _libs["dropin"] = lib
        return _libs["dropin"]


def dropin_model_wave(model, sb, x):
    with _lock:
        _ref_init(model)
        y = np.zeros(sb.dim)
        dropin_lib().dropin_model_wave(sb.sys_len, sb.env_len, *_m(sb.sys), *_m(sb.env), *_bs(sb.sys_bs),
                                       *_bs(sb.env_bs), int(sb.wsec.shape[0]), np.ascontiguousarray(sb.wsec),
                                       np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def dropin_lanczos(model, sb, x0, lanzero=1e-10):
    with _lock:
        _ref_init(model, lanzero)
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        e0 = C.c_double(0.0)
        dropin_lib().dropin_lanczos(sb.sys_len, sb.env_len, *_m(sb.sys), *_m(sb.env), *_bs(sb.sys_bs),
                                    *_bs(sb.env_bs), int(sb.wsec.shape[0]), np.ascontiguousarray(sb.wsec), x,
                                    C.byref(e0))
    return e0.value, x
