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


class _quiet:
    """Sends what the reference prints on its standard output (unit 6) to
    standard error, so a caller's standard output stays its own."""

    def __enter__(self):
        import sys
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        lib = _libs.get("ref")
        if lib is not None:
            lib.refb_flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)

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
            lib.oracle_opset_sizes.restype = C.c_int
            lib.oracle_opset_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int64)]
            lib.oracle_opset_copy.restype = None
            lib.oracle_opset_copy.argtypes = [C.c_void_p, I32P, I32P, I64P, F64P]
            lib.oracle_opset_free.restype = None
            lib.oracle_opset_free.argtypes = [C.c_void_p]
            lib.oracle_get_hamiltonian.restype = C.c_void_p
            lib.oracle_get_hamiltonian.argtypes = ([C.POINTER(Lattice), C.c_int, C.c_int, C.c_int] + _model_args()
                                                   + [C.c_int, C.c_int, I32P, I32P, I32P, I32P, I64P, F64P])
            lib.oracle_update_trun.restype = C.c_void_p
            lib.oracle_update_trun.argtypes = _model_args() + [I32P, I32P, I32P, I64P, F64P]
            lib.oracle_density_matrix.restype = C.c_void_p
            lib.oracle_density_matrix.argtypes = [C.c_int, C.c_int, C.c_int, I32P, C.c_int, F64P]
            lib.oracle_quicksort_double.restype = None
            lib.oracle_quicksort_double.argtypes = [F64P, I32P, C.c_int]
            lib.oracle_wave_trans.restype = C.c_int
            lib.oracle_wave_trans.argtypes = ([C.c_int, C.c_int, I32P, I32P, C.c_int, I32P, I32P]
                                              + [I32P, I32P, I64P, F64P] * 2
                                              + [C.c_int, I32P, F64P, C.c_int, I32P, F64P])
            lib.oracle_truncation_select.restype = C.c_int
            lib.oracle_truncation_select.argtypes = [C.c_int, I32P, F64P, C.c_int, C.c_int, C.c_double, I32P, F64P,
                                                     F64P, I32P]
            _libs["oracle"] = lib
        return _libs["oracle"]


def ref_path() -> str:
    return os.path.join(HERE, "_ref", "libtjsq_ref.so")


def ref_available() -> bool:
    return os.path.exists(ref_path())


def ref_threads(n: int) -> None:
    """Threads MKL uses inside the reference's DGEMM / DSYEV calls (the
    reference builds with -qmkl=parallel, code/tjsq/makefile:3)."""
    ref_lib().MKL_Set_Num_Threads(int(n))


def ref_lib():
    with _lock:
        if "ref" not in _libs:
            lib = C.CDLL(ref_path(), mode=C.RTLD_GLOBAL)
            # MKL through mkl_rt: one thread unless a caller asks for more
            # (parity runs stay deterministic)
            lib.MKL_Set_Num_Threads.restype = None
            lib.MKL_Set_Num_Threads.argtypes = [C.c_int]
            lib.MKL_Set_Num_Threads(1)
            lib.refb_init.restype = C.c_int
            lib.refb_init.argtypes = [C.c_int, C.c_int] + [C.c_double] * 8
            margs = [C.c_int, I32P, I32P, I64P, F64P]
            common = [C.c_int, C.c_int] + margs + margs + [C.c_int, I32P, I32P, C.c_int, I32P, I32P, C.c_int, I32P]
            lib.refb_model_wave.restype = C.c_int
            lib.refb_model_wave.argtypes = common + [F64P, F64P]
            lib.refb_lanczos.restype = C.c_int
            lib.refb_lanczos.argtypes = common + [F64P, C.POINTER(C.c_double)]
            lib.refb_get_super_basis.restype = C.c_int
            lib.refb_get_super_basis.argtypes = ([C.c_int, C.c_int, I32P, I32P] * 2
                                                 + [C.c_int, C.c_int, I32P, C.POINTER(C.c_int), C.POINTER(C.c_int)])
            lib.refb_w6js.restype = C.c_double
            lib.refb_w6js.argtypes = [C.c_int] * 6
            lib.refb_get_basis.restype = C.c_int
            lib.refb_get_basis.argtypes = [C.c_int, C.c_int, I32P, C.c_int, I32P, I32P, I32P, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
            lib.refb_result_sizes.restype = C.c_int
            lib.refb_result_sizes.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]
            lib.refb_result_copy.restype = C.c_int
            lib.refb_result_copy.argtypes = [I32P, I32P, I64P, F64P]
            lib.refb_get_hamiltonian.restype = C.c_int
            lib.refb_get_hamiltonian.argtypes = ([C.c_int, C.c_int, C.c_int] + margs
                                                 + [C.c_int, C.c_int, I32P, I32P, I32P, I32P, I64P, F64P])
            lib.refb_update_trun.restype = C.c_int
            lib.refb_update_trun.argtypes = margs + [I32P, I32P, I32P, I64P, F64P]
            lib.refb_density.restype = C.c_int
            lib.refb_density.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, I32P, C.c_int, F64P]
            lib.refb_truncation.restype = C.c_int
            lib.refb_truncation.argtypes = [I32P, I32P, I64P, F64P, C.c_int, C.c_int, C.c_double, C.c_int,
                                            C.POINTER(C.c_int), C.POINTER(C.c_double)]
            lib.refb_quicksort.restype = C.c_int
            lib.refb_quicksort.argtypes = [C.c_int, F64P, I32P]
            lib.refb_wave_trans.restype = C.c_int
            lib.refb_wave_trans.argtypes = ([C.c_int, C.c_int, I32P, I32P, C.c_int, I32P, I32P]
                                            + [I32P, I32P, I64P, F64P] * 2
                                            + [C.c_int, I32P, F64P, C.c_int, I32P, F64P])
            lib.refb_flush.restype = C.c_int
            lib.refb_flush.argtypes = []
            lib.refb_result_dims.restype = C.c_int
            lib.refb_result_dims.argtypes = [I32P]
            lib.refb_model_to_disk.restype = C.c_int
            lib.refb_model_to_disk.argtypes = [C.c_int] + margs + [I32P, C.c_int, C.c_int]
            lib.refb_model_from_disk.restype = C.c_int
            lib.refb_model_from_disk.argtypes = [C.c_int, C.c_int]
            lib.refb_truns_to_disk.restype = C.c_int
            lib.refb_truns_to_disk.argtypes = [I32P, I32P, I64P, F64P, C.c_int, C.c_int, C.c_int]
            lib.refb_truns_from_disk.restype = C.c_int
            lib.refb_truns_from_disk.argtypes = [C.c_int, C.c_int]
            lib.refb_basis_to_disk.restype = C.c_int
            lib.refb_basis_to_disk.argtypes = [C.c_int, C.c_int, I32P, I32P, C.c_int, C.c_int]
            lib.refb_basis_from_disk.restype = C.c_int
            lib.refb_basis_from_disk.argtypes = [C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 3
            lib.refb_basis_copy.restype = C.c_int
            lib.refb_basis_copy.argtypes = [I32P, I32P, I32P]
            lib.refb_wave_to_disk.restype = C.c_int
            lib.refb_wave_to_disk.argtypes = [C.c_int, C.c_int, C.c_int, I32P, F64P, C.c_char_p, C.c_int]
            lib.refb_wave_from_disk.restype = C.c_int
            lib.refb_wave_from_disk.argtypes = [C.c_char_p, C.c_int] + [C.POINTER(C.c_int)] * 4
            lib.refb_wave_copy.restype = C.c_int
            lib.refb_wave_copy.argtypes = [I32P, F64P]
            lib.refb_set_bsm.restype = C.c_int
            lib.refb_set_bsm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, I32P, I32P]
            lib.refb_set_truns.restype = C.c_int
            lib.refb_set_truns.argtypes = [C.c_int, C.c_int, I32P, I32P, I64P, F64P]
            lib.refb_oper_cor_dia.restype = C.c_int
            lib.refb_oper_cor_dia.argtypes = [C.c_int, C.c_int, C.c_int, I32P, F64P, F64P]
            lib.refb_cross_cor.restype = C.c_int
            lib.refb_cross_cor.argtypes = [C.c_int, C.c_int, C.c_int, I32P, F64P, F64P]
            lib.refb_measure_density.restype = C.c_int
            lib.refb_measure_density.argtypes = [C.c_int, C.c_int, C.c_int, I32P, F64P, C.c_int, F64P]
            lib.refb_w3js.restype = C.c_double
            lib.refb_w3js.argtypes = [C.c_int] * 6
            lib.refb_density_cor.restype = C.c_int
            lib.refb_density_cor.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, I32P, F64P, F64P]
            lib.refb_sc_order.restype = C.c_int
            lib.refb_sc_order.argtypes = [C.c_int] * 6 + [I32P, C.c_int, I32P, F64P, F64P]
            lib.refb_sc_pair_cor0.restype = C.c_int
            lib.refb_sc_pair_cor0.argtypes = [C.c_int] * 6 + [I32P, C.c_int, I32P, F64P, C.c_int, F64P]
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
    with _quiet():
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
        with _quiet():
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
        with _quiet():
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
            with _quiet():
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
        lib = ref_lib()
        with _quiet():
            lib.refb_get_basis(block_len, blk.shape[0], blk, site.shape[0], site, sec, sub, C.byref(ns),
                               C.byref(nq))
    return sec[:4 * ns.value].reshape(-1, 4), sub[:6 * nq.value].reshape(-1, 6)


def oracle_get_super_basis(sys_bs, env_bs, total_up, total_down):
    """The C restatement's Get_super_basis: the wavefunction sector table."""
    ssec = np.ascontiguousarray(sys_bs.sec, dtype=np.int32)
    esec = np.ascontiguousarray(env_bs.sec, dtype=np.int32)
    w = np.zeros(6 * max(1, ssec.shape[0] * esec.shape[0]), np.int32)
    n = oracle_lib().oracle_get_super_basis(ssec.shape[0], ssec, esec.shape[0], esec, int(total_up),
                                            int(total_down), w)
    return w[:6 * n].reshape(-1, 6)


def ref_get_super_basis(sys_bs, env_bs, total_up, total_down):
    """Get_super_basis of the reference build (block.f90:110-161): the
    wavefunction sector table (sys_up, sys_down, env_up, env_down, sys_dim,
    env_dim) in the routine's order, and the routine's total dimension."""
    ssec = np.ascontiguousarray(sys_bs.sec, dtype=np.int32)
    ssub = np.ascontiguousarray(sys_bs.sub, dtype=np.int32)
    esec = np.ascontiguousarray(env_bs.sec, dtype=np.int32)
    esub = np.ascontiguousarray(env_bs.sub, dtype=np.int32)
    w = np.zeros(6 * max(1, ssec.shape[0] * esec.shape[0]), np.int32)
    nw, dim = C.c_int(0), C.c_int(0)
    with _lock:
        lib = ref_lib()
        with _quiet():
            lib.refb_get_super_basis(sys_bs.length, ssec.shape[0], ssec, ssub, env_bs.length, esec.shape[0], esec,
                                     esub, int(total_up), int(total_down), w, C.byref(nw), C.byref(dim))
    return w[:6 * nw.value].reshape(-1, 6), dim.value


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
            lib.MKL_Set_Num_Threads.restype = None
            lib.MKL_Set_Num_Threads.argtypes = [C.c_int]
            lib.MKL_Set_Num_Threads(1)     # the reference routines inside it, as in ref_lib
            margs = [C.c_int, I32P, I32P, I64P, F64P]
            common = [C.c_int, C.c_int] + margs + margs + [C.c_int, I32P, I32P, C.c_int, I32P, I32P, C.c_int, I32P]
            lib.dropin_model_wave.restype = C.c_int
            lib.dropin_model_wave.argtypes = common + [F64P, F64P]
            lib.dropin_measure_density.restype = C.c_int
            lib.dropin_measure_density.argtypes = [C.c_int, C.c_int, C.c_int, I32P, F64P, C.c_int, F64P]
            lib.dropin_lanczos.restype = C.c_int
            lib.dropin_lanczos.argtypes = common + [F64P, C.POINTER(C.c_double)]
            ref = ref_lib()
            # the drop-in test entry points take the reference bridge's
            # arguments (dropin_truncation without system_block)
            for name in ("density", "truncation", "get_hamiltonian", "wave_trans"):
                getattr(lib, "dropin_" + name).restype = C.c_int
            lib.dropin_density.argtypes = ref.refb_density.argtypes
            lib.dropin_get_hamiltonian.argtypes = ref.refb_get_hamiltonian.argtypes
            lib.dropin_wave_trans.argtypes = ref.refb_wave_trans.argtypes
            lib.dropin_truncation.argtypes = ref.refb_truncation.argtypes
            lib.dropin_plan_builds.restype = C.c_int
            lib.dropin_plan_builds.argtypes = []
            lib.dropin_release_plan.restype = C.c_int
            lib.dropin_release_plan.argtypes = []
            lib.dropin_set_jt.restype = C.c_int
            lib.dropin_set_jt.argtypes = [C.c_int, C.c_int, C.c_double]
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
        with _quiet():
            dropin_lib().dropin_lanczos(sb.sys_len, sb.env_len, *_m(sb.sys), *_m(sb.env), *_bs(sb.sys_bs),
                                        *_bs(sb.env_bs), int(sb.wsec.shape[0]), np.ascontiguousarray(sb.wsec), x,
                                        C.byref(e0))
    return e0.value, x


# ---------------------------------------------------------------------------
# block update: density matrix, truncation, enlargement and projection
# ---------------------------------------------------------------------------
def _opset(length, hdr, sub, off, data):
    from types import SimpleNamespace
    return SimpleNamespace(length=int(length), nops=int(hdr.shape[0]), hdr=hdr, sub=sub, off=off, data=data)


def _take_oracle(handle, length=None):
    lib = oracle_lib()
    if not handle:
        raise RuntimeError("oracle call failed (inconsistent tables)")
    no, ns, nd = C.c_int(0), C.c_int(0), C.c_int64(0)
    lib.oracle_opset_sizes(handle, C.byref(no), C.byref(ns), C.byref(nd))
    hdr = np.zeros((no.value, 4), np.int32)
    sub = np.zeros((ns.value, 5), np.int32)
    off = np.zeros(ns.value, np.int64)
    data = np.zeros(max(1, nd.value), np.float64)
    lib.oracle_opset_copy(handle, hdr, sub, off, data)
    lib.oracle_opset_free(handle)
    return _opset(hdr[0, 0] if length is None else length, hdr, sub, off, data[:nd.value])


def _take_ref(length=None):
    lib = ref_lib()
    no, ns, nd = C.c_int(0), C.c_int(0), C.c_int64(0)
    lib.refb_result_sizes(C.byref(no), C.byref(ns), C.byref(nd))
    hdr = np.zeros((no.value, 4), np.int32)
    sub = np.zeros((ns.value, 5), np.int32)
    off = np.zeros(ns.value, np.int64)
    data = np.zeros(max(1, nd.value), np.float64)
    lib.refb_result_copy(hdr, sub, off, data)
    return _opset(hdr[0, 0] if length is None else length, hdr, sub, off, data[:nd.value])


def _one(op):
    """Flat arrays of a one-operator set (truncation operators)."""
    return [np.ascontiguousarray(op.hdr[0], np.int32), np.ascontiguousarray(op.sub, np.int32),
            np.ascontiguousarray(op.off, np.int64), np.ascontiguousarray(op.data, np.float64)]


def get_hamiltonian(model, side, sys_len, env_len, pre, basis, trun):
    """Oracle Get_Hamiltonian_Sys (side 0) / _Env (side 1): the new block model."""
    lib = oracle_lib()
    lat = lattice_struct(model)
    h = lib.oracle_get_hamiltonian(C.byref(lat), side, sys_len, env_len, *_m(pre), basis.length, *_bs(basis),
                                   *_one(trun))
    return _take_oracle(h)


def update_trun(ops, ndia, trun):
    """Oracle update_trun_dia / update_trun_ndia of every operator."""
    h = oracle_lib().oracle_update_trun(*_m(ops), np.ascontiguousarray(ndia, np.int32), *_one(trun))
    return _take_oracle(h)


def density_matrix(side, length, wsec, psi, levels=1):
    """Oracle Get_density_matrix_sys/env_new."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    h = oracle_lib().oracle_density_matrix(side, length, w.shape[0], w, levels,
                                          np.ascontiguousarray(psi, np.float64))
    return _take_oracle(h, length)


def _wave_trans_args(direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, x, wsec_new):
    wo = np.ascontiguousarray(wsec_old, np.int32).reshape(-1, 6)
    wn = np.ascontiguousarray(wsec_new, np.int32).reshape(-1, 6)
    dn = int((wn[:, 4].astype(np.int64) * wn[:, 5]).sum())
    y = np.zeros(max(dn, 1))
    args = ([direction] + _bs(sys_bs) + _bs(env_bs) + _one(sys_trun) + _one(env_trun)
            + [wo.shape[0], wo, np.ascontiguousarray(x, np.float64), wn.shape[0], wn, y])
    return args, y[:dn]


def wave_trans(direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, x, wsec_new):
    """Oracle wave_trans_LR (direction 0) / wave_trans_RL (1) + wave_normalize."""
    args, y = _wave_trans_args(direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, x, wsec_new)
    if oracle_lib().oracle_wave_trans(*args) != 0:
        raise ValueError("truncation operator dimensions disagree with the wavefunction sectors")
    return y


def quicksort_double(arr):
    """Oracle QuickSort_Double: (values, 1-based permutation)."""
    a = np.ascontiguousarray(arr, np.float64).copy()
    idx = np.arange(1, a.shape[0] + 1, dtype=np.int32)
    oracle_lib().oracle_quicksort_double(a, idx, a.shape[0])
    return a, idx


def truncation_select(den_sub, eigval, kept_min, kept_max, trun_err):
    """Oracle selection of Get_truncation_operator on given eigenvalues:
    (counts, info dict, sorted spectrum, order)."""
    den_sub = np.ascontiguousarray(den_sub, np.int32).reshape(-1, 5)
    ev = np.ascontiguousarray(eigval, np.float64)
    counts = np.zeros(den_sub.shape[0], np.int32)
    info = np.zeros(5, np.float64)
    srt = np.zeros(ev.shape[0], np.float64)
    order = np.zeros(ev.shape[0], np.int32)
    oracle_lib().oracle_truncation_select(den_sub.shape[0], den_sub, ev, kept_min, kept_max, trun_err, counts, info,
                                          srt, order)
    d = {"num_state": int(info[0]), "trunc_err": info[1], "entropy": info[2], "entropy1": info[3],
         "entropy2": info[4]}
    return counts, d, srt, order


def truncation_operator(den, kept_min, kept_max, trun_err):
    """Oracle Get_truncation_operator: per-sector symmetric eigenproblems by
    numpy.linalg.eigh (LAPACK syevd; the reference calls DSYEV on the same
    matrices, den_mat.f90:192), then the restated selection.  Returns
    (trun, counts, info, spectrum, eigvals per sector)."""
    subs = np.asarray(den.sub).reshape(-1, 5)
    evs, vecs = [], []
    for s in range(subs.shape[0]):
        n = int(subs[s, 4])
        o = int(den.off[s])
        m = np.asarray(den.data[o:o + n * n]).reshape(n, n).T
        w, v = np.linalg.eigh(m)
        evs.append(w)
        vecs.append(v)
    ev = np.concatenate(evs) if evs else np.zeros(0)
    counts, info, spec, _ = truncation_select(subs, ev, kept_min, kept_max, trun_err)
    tsub, toff, chunks, nd = [], [], [], 0
    for s in range(subs.shape[0]):
        c = int(counts[s])
        if c <= 0:
            continue
        n = int(subs[s, 4])
        tsub.append((subs[s, 0], subs[s, 1], 0, n, c))
        toff.append(nd)
        chunks.append(np.asarray(vecs[s][:, n - c:], np.float64).T.reshape(-1))
        nd += n * c
    trun = _opset(den.length, np.array([[den.length, 0, 0, len(tsub)]], np.int32),
                  np.array(tsub, np.int32).reshape(-1, 5), np.array(toff, np.int64),
                  np.concatenate(chunks) if chunks else np.zeros(0))
    return trun, counts, info, spec, evs


def ref_get_hamiltonian(model, side, sys_len, env_len, pre, basis, trun):
    """Get_Hamiltonian_Sys / _Env of the reference build."""
    with _lock:
        lib = _ref_init(model)
        lib.refb_get_hamiltonian(side, sys_len, env_len, *_m(pre), basis.length, *_bs(basis), *_one(trun))
        return _take_ref()


def ref_update_trun(ops, ndia, trun):
    with _lock:
        lib = ref_lib()
        lib.refb_update_trun(*_m(ops), np.ascontiguousarray(ndia, np.int32), *_one(trun))
        return _take_ref()


def ref_wave_trans(direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, x, wsec_new):
    """wave_trans_LR / wave_trans_RL of the reference build."""
    args, y = _wave_trans_args(direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, x, wsec_new)
    with _lock:
        ref_lib().refb_wave_trans(*args)
    return y


def ref_density_matrix(side, sys_len, env_len, wsec, psi, levels=1):
    """Get_density_matrix_sys/env_new of the reference build (wave lengths
    sys_len / env_len are the enlarged block lengths)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    with _lock:
        lib = ref_lib()
        lib.refb_density(side, sys_len, env_len, w.shape[0], w, levels, np.ascontiguousarray(psi, np.float64))
        return _take_ref()


def ref_truncation(den, kept_min, kept_max, trun_err, workdir, system_block=True):
    """Get_truncation_operator of the reference build, run with the working
    directory set to ``workdir`` (it appends to three log files there).
    Returns (trun, num_state, Entropys)."""
    ns, ent = C.c_int(0), C.c_double(0.0)
    with _lock:
        lib = ref_lib()
        cwd = os.getcwd()
        os.chdir(workdir)
        try:
            with _quiet():
                lib.refb_truncation(*_one(den), kept_min, kept_max, trun_err, int(system_block), C.byref(ns),
                                    C.byref(ent))
        finally:
            os.chdir(cwd)
        return _take_ref(), ns.value, ent.value


def dropin_density_matrix(side, sys_len, env_len, wsec, psi, levels=1):
    """Get_density_matrix_sys/env_new through the product's Fortran drop-in
    (as ref_density_matrix)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    with _lock:
        dropin_lib().dropin_density(side, sys_len, env_len, w.shape[0], w, levels, np.ascontiguousarray(psi, np.float64))
        return _take_ref()


def dropin_truncation(den, kept_min, kept_max, trun_err, workdir=None, system_block=True):
    """Get_truncation_operator through the drop-in (as ref_truncation; its
    log files go to ``workdir``, default the current directory)."""
    ns, ent = C.c_int(0), C.c_double(0.0)
    with _lock:
        cwd = os.getcwd()
        if workdir:
            os.chdir(workdir)
        try:
            dropin_lib().dropin_truncation(*_one(den), kept_min, kept_max, trun_err, int(system_block), C.byref(ns),
                                           C.byref(ent))
        finally:
            os.chdir(cwd)
        return _take_ref(), ns.value, ent.value


def dropin_plan_builds() -> int:
    """Plans the drop-in has compiled so far (its plan cache's misses)."""
    with _lock:
        return int(dropin_lib().dropin_plan_builds())


def dropin_release_plan():
    with _lock:
        dropin_lib().dropin_release_plan()


def dropin_set_jt(k, i, v):
    """jt(k, i) of the reference's pubdata (1-based direction and site)."""
    with _lock:
        dropin_lib().dropin_set_jt(k, i, v)


def dropin_get_hamiltonian(model, side, sys_len, env_len, pre, basis, trun):
    """Get_Hamiltonian_Sys / _Env through the drop-in (as ref_get_hamiltonian)."""
    with _lock:
        _ref_init(model)
        dropin_lib().dropin_get_hamiltonian(side, sys_len, env_len, *_m(pre), basis.length, *_bs(basis), *_one(trun))
        return _take_ref()


def dropin_wave_trans(direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, x, wsec_new):
    """wave_trans_LR / wave_trans_RL through the drop-in (as ref_wave_trans)."""
    args, y = _wave_trans_args(direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, x, wsec_new)
    with _lock:
        dropin_lib().dropin_wave_trans(*args)
    return y


# ---------------------------------------------------------------------------
# checkpoint files: the reference's own writers and readers (block.f90:2760-3097)
# ---------------------------------------------------------------------------
class _cwd:
    def __init__(self, path):
        self.path = os.fspath(path)

    def __enter__(self):
        self.saved = os.getcwd()
        os.chdir(self.path)

    def __exit__(self, *exc):
        os.chdir(self.saved)


def _result_dims(n):
    d = np.zeros(max(1, n), np.int32)
    ref_lib().refb_result_dims(d)
    return d[:n].astype(np.int64)


def ref_result_dims():
    """The `dim` fields of the operators the last reference call produced."""
    with _lock:
        no, ns, nd = C.c_int(0), C.c_int(0), C.c_int64(0)
        ref_lib().refb_result_sizes(C.byref(no), C.byref(ns), C.byref(nd))
        return _result_dims(no.value)


def ref_model_to_disk(model, workdir, ops, dims, idx, system=True):
    """model_to_disk of the reference build into workdir (the model of a
    lattice like ``model``'s: 1 + 4 nleg11(len) operators)."""
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        rc = lib.refb_model_to_disk(ops.length, *_m(ops), np.ascontiguousarray(dims, np.int32), idx, int(system))
    if rc != 0:
        raise RuntimeError("model does not hold 1 + 4 nleg11(len) operators")


def ref_model_from_disk(model, workdir, idx, system=True):
    """model_from_disk of the reference build: (ops, dims)."""
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        length = lib.refb_model_from_disk(idx, int(system))
        ops = _take_ref(length)
        return ops, _result_dims(ops.nops)


def ref_truns_to_disk(workdir, trun, dim, idx, system=True):
    with _lock, _cwd(workdir):
        ref_lib().refb_truns_to_disk(*_one(trun), int(dim), idx, int(system))


def ref_truns_from_disk(workdir, idx, system=True):
    with _lock, _cwd(workdir):
        ref_lib().refb_truns_from_disk(idx, int(system))
        t = _take_ref()
        return t, int(_result_dims(1)[0])


def ref_basis_to_disk(workdir, basis, idx, system=True):
    with _lock, _cwd(workdir):
        ref_lib().refb_basis_to_disk(basis.length, *_bs(basis), idx, int(system))


def ref_basis_from_disk(workdir, idx, system=True):
    """basis_from_disk of the reference build: (length, sec, sub, idx)."""
    ln, nsec, nsub = C.c_int(0), C.c_int(0), C.c_int(0)
    with _lock, _cwd(workdir):
        lib = ref_lib()
        lib.refb_basis_from_disk(idx, int(system), C.byref(ln), C.byref(nsec), C.byref(nsub))
        sec = np.zeros((max(1, nsec.value), 4), np.int32)
        sub = np.zeros((max(1, nsub.value), 6), np.int32)
        bidx = np.zeros(max(1, nsec.value), np.int32)
        lib.refb_basis_copy(sec, sub, bidx)
    return ln.value, sec[:nsec.value], sub[:nsub.value], bidx[:nsec.value]


def ref_wave_to_disk(workdir, name, sys_len, env_len, wsec, x):
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    with _lock, _cwd(workdir):
        ref_lib().refb_wave_to_disk(sys_len, env_len, w.shape[0], w, np.ascontiguousarray(x, np.float64),
                                    name.encode(), len(name))


def ref_wave_from_disk(workdir, name):
    """wave_from_disk of the reference build: (sys_len, env_len, wsec, x)."""
    sl, el, nw, dim = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    with _lock, _cwd(workdir):
        lib = ref_lib()
        lib.refb_wave_from_disk(name.encode(), len(name), C.byref(sl), C.byref(el), C.byref(nw), C.byref(dim))
        wsec = np.zeros((max(1, nw.value), 6), np.int32)
        y = np.zeros(max(1, dim.value), np.float64)
        lib.refb_wave_copy(wsec, y)
    return sl.value, el.value, wsec[:nw.value], y[:dim.value]


def ref_measure_density(model, workdir, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, x):
    """measure_operator_dia of num_elec in the reference build (it appends
    to num_elec.dat in workdir): <n_i> of the sys_len + env_len sites of the
    wavefunction, in lattice order.  Bases keyed by length 2..sys_len
    (env_len), truncation operators by 2..sys_len - 1 (env_len - 1)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        for side, bases, truns in ((0, sys_bases, sys_truns), (1, env_bases, env_truns)):
            for L, b in bases.items():
                lib.refb_set_bsm(side, L, b.length, *_bs(b))
            for L, t in truns.items():
                lib.refb_set_truns(side, L, *_one(t))
        ave = np.zeros(4096)
        with _quiet():
            nref = lib.refb_measure_density(sys_len, env_len, w.shape[0], w, np.ascontiguousarray(x, np.float64), 1,
                                            ave)
    n = sys_len + env_len
    out = np.zeros(n)
    out[:sys_len] = ave[:sys_len]
    for i in range(1, env_len + 1):
        out[n - i] = ave[nref - i]
    return out


def ref_cross_correlation(model, workdir, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, x):
    """The across-the-bond pairs of Get_oper_cor_dia for num_elec, computed
    with the reference's own transports and cor_dia routines: a
    (sys_len, env_len) matrix, rows system-side sites 1..sys_len, columns
    environment-side sites 1..env_len (the last of each the bond site)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        for side, bases, truns in ((0, sys_bases, sys_truns), (1, env_bases, env_truns)):
            for L, b in bases.items():
                lib.refb_set_bsm(side, L, b.length, *_bs(b))
            for L, t in truns.items():
                lib.refb_set_truns(side, L, *_one(t))
        cor = np.zeros(sys_len * env_len)
        lib.refb_cross_cor(sys_len, env_len, w.shape[0], w, np.ascontiguousarray(x, np.float64), cor)
    return cor.reshape(env_len, sys_len).T


DENSITY_COR_KINDS = {"density": 0, "spin": 1, "hopping": 2}


def ref_w3js(*j) -> float:
    """W3JS of the reference build."""
    global _ref_ready
    with _lock:
        lib = ref_lib()
        if not _ref_ready:
            with _quiet():
                lib.refb_init(6, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1e-10)
            _ref_ready = True
        return lib.refb_w3js(*[int(v) for v in j])


def ref_density_cor(model, workdir, kind, sid1, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec,
                    x):
    """The pairs of Get_density_cor (measure.f90:3110-3223) of system-side
    site sid1 with every later site of the chain, computed with the
    reference's own transports, products and cor_ndia routines (kind
    "density": num_elec x num_elec, "spin": st_sd x st_sd, "hopping":
    st_elec_up x st_elec_down with fermion signs).  Returns cor (length
    sys_len + env_len, lattice order, entry sid - 1; NaN for sid <= sid1)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    n = sys_len + env_len
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        for side, bases, truns in ((0, sys_bases, sys_truns), (1, env_bases, env_truns)):
            for L, b in bases.items():
                lib.refb_set_bsm(side, L, b.length, *_bs(b))
            for L, t in truns.items():
                lib.refb_set_truns(side, L, *_one(t))
        cor = np.full(n, np.nan)
        rc = lib.refb_density_cor(DENSITY_COR_KINDS[kind], sid1, sys_len, env_len, w.shape[0], w,
                                  np.ascontiguousarray(x, np.float64), cor)
    if rc != n:
        raise RuntimeError(f"refb_density_cor returned {rc}")
    cor[:sid1] = np.nan
    return cor


def _ref_chain(lib, sys_bases, env_bases, sys_truns, env_truns):
    for side, bases, truns in ((0, sys_bases, sys_truns), (1, env_bases, env_truns)):
        for L, b in bases.items():
            lib.refb_set_bsm(side, L, b.length, *_bs(b))
        for L, t in truns.items():
            lib.refb_set_truns(side, L, *_one(t))


def ref_sc_order(model, workdir, neib, ny, xc1, xc2, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns,
                 wsec, x):
    """Get_SC_Order (measure.f90:3226-3376) over a chain of sys_len +
    env_len sites with neighbour table neib (neib[i, k - 1] = neib(i, k)),
    computed with the reference's own routines (refb_sc_order).  Returns
    rows (sid1, neib[sid1, 2], val1(1), val1(2))."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    n = sys_len + env_len
    nb = np.ascontiguousarray(np.asarray(neib, np.int32)[:n + 1, :12])
    sids = list(range(xc1, xc2 + 1, ny))
    out = np.zeros(2 * max(1, len(sids)))
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        _ref_chain(lib, sys_bases, env_bases, sys_truns, env_truns)
        with _quiet():
            rc = lib.refb_sc_order(xc1, xc2, ny, sys_len, env_len, n, nb, w.shape[0], w,
                                   np.ascontiguousarray(x, np.float64), out)
    if rc != len(sids):
        raise RuntimeError(f"refb_sc_order returned {rc}")
    return [(s, int(nb[s, 2]), out[2 * k], out[2 * k + 1]) for k, s in enumerate(sids)]


def ref_sc_pair_cor0(model, workdir, neib, ny, xc1, ij, sys_len, env_len, sys_bases, env_bases, sys_truns,
                     env_truns, wsec, x, maxrows=256):
    """Get_mSCOP_Operator_Cor0 (measure.f90:3377-3507) with xc1 = xc2 over a
    chain of sys_len + env_len sites, computed with the reference's own
    routines (refb_sc_pair_cor0): rows (sid1, sid3, mod, dx, v1, v3, v5)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    n = sys_len + env_len
    nb = np.ascontiguousarray(np.asarray(neib, np.int32)[:n + 1, :12])
    out = np.zeros((maxrows, 7))
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        _ref_chain(lib, sys_bases, env_bases, sys_truns, env_truns)
        with _quiet():
            rc = lib.refb_sc_pair_cor0(xc1, ij, ny, sys_len, env_len, n, nb, w.shape[0], w,
                                       np.ascontiguousarray(x, np.float64), maxrows, out)
    if rc < 0 or rc > maxrows:
        raise RuntimeError(f"refb_sc_pair_cor0 returned {rc}")
    return [tuple(int(v) for v in r[:4]) + tuple(float(v) for v in r[4:]) for r in out[:rc]]


def ref_oper_cor_dia(model, workdir, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, x):
    """The reference's whole Get_oper_cor_dia (measure.f90:468-634) of
    num_elec: the (n, n) matrix oper_dia over the wavefunction's
    sys_len + env_len sites in lattice order (it appends to two files in
    workdir)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        for side, bases, truns in ((0, sys_bases, sys_truns), (1, env_bases, env_truns)):
            for L, b in bases.items():
                lib.refb_set_bsm(side, L, b.length, *_bs(b))
            for L, t in truns.items():
                lib.refb_set_truns(side, L, *_one(t))
        cor = np.zeros(4096 * 4096 // 64)
        with _quiet():
            nref = lib.refb_oper_cor_dia(sys_len, env_len, w.shape[0], w, np.ascontiguousarray(x, np.float64), cor)
    full = cor[:nref * nref].reshape(nref, nref).T
    idx = list(range(sys_len)) + [nref - j for j in range(env_len, 0, -1)]
    return full[np.ix_(idx, idx)]


def dropin_measure_density(model, workdir, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, x,
                           trun_idx=1):
    """measure_operator_dia through the product's Fortran drop-in (as
    ref_measure_density, whose globals the drop-in shares)."""
    w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
    with _lock, _cwd(workdir):
        lib = _ref_init(model)
        for side, bases, truns in ((0, sys_bases, sys_truns), (1, env_bases, env_truns)):
            for L, b in bases.items():
                lib.refb_set_bsm(side, L, b.length, *_bs(b))
            for L, t in truns.items():
                lib.refb_set_truns(side, L, *_one(t))
        ave = np.zeros(4096)
        with _quiet():
            nref = dropin_lib().dropin_measure_density(sys_len, env_len, w.shape[0], w,
                                                      np.ascontiguousarray(x, np.float64), trun_idx, ave)
    n = sys_len + env_len
    out = np.zeros(n)
    out[:sys_len] = ave[:sys_len]
    for i in range(1, env_len + 1):
        out[n - i] = ave[nref - i]
    return out


def ref_quicksort(arr):
    a = np.ascontiguousarray(arr, np.float64).copy()
    idx = np.arange(1, a.shape[0] + 1, dtype=np.int32)
    with _lock:
        ref_lib().refb_quicksort(a.shape[0], a, idx)
    return a, idx
