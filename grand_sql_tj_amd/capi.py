"""ctypes binding of libtjsq_hip.so (the C-ABI declared in include/tjsq_capi.h).

The library is built in-tree by ``grand_sql_tj_amd.build.build()``.  Loading
fails loudly when it is missing: there is no CPU fallback for the apply or the
solver.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtjsq_hip.so")

TJSQ_OK = 0
TJSQ_ERR_ARG = -1
FP64, FP32, BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}


class Lattice(C.Structure):
    _fields_ = [("ny", C.c_int), ("nx_tables", C.c_int), ("jt", C.c_double * 2), ("jd", C.c_double * 2),
                ("jn", C.c_double * 2), ("esite", C.c_double), ("up_bias", C.c_int)]


class OpSetC(C.Structure):
    _fields_ = [("length", C.c_int), ("nops", C.c_int), ("hdr", C.c_void_p), ("sub", C.c_void_p),
                ("off", C.c_void_p), ("data", C.c_void_p), ("ndata", C.c_int64), ("data_on_device", C.c_int)]


class BasisC(C.Structure):
    _fields_ = [("length", C.c_int), ("nsec", C.c_int), ("sec", C.c_void_p), ("sub", C.c_void_p)]


class PlanStats(C.Structure):
    _fields_ = [("dim", C.c_int64), ("n_contrib", C.c_int64), ("n_chain", C.c_int64), ("n_stage1", C.c_int64),
                ("n_out_blocks", C.c_int64), ("n_tiles1", C.c_int64), ("n_tiles2", C.c_int64),
                ("flops_ref", C.c_double), ("flops_exec", C.c_double), ("bytes_fp64", C.c_double),
                ("op_elems", C.c_int64), ("scratch_elems", C.c_int64), ("n_slots", C.c_int64),
                ("packed", C.c_int64), ("flops_packed", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TruncInfo(C.Structure):
    _fields_ = [("num_state", C.c_int), ("dim", C.c_int), ("trunc_err", C.c_double), ("entropy", C.c_double),
                ("entropy1", C.c_double), ("entropy2", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LanczosOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("precision", C.c_int), ("refine_tol", C.c_double),
                ("max_refine", C.c_int), ("energy_tol", C.c_double), ("rtol", C.c_double)]


class LanczosResult(C.Structure):
    _fields_ = [("e0", C.c_double), ("e_ritz", C.c_double), ("residual", C.c_double), ("iterations", C.c_int),
                ("applies", C.c_int), ("refinements", C.c_int), ("breakdown", C.c_int), ("refine_iterations", C.c_int),
                ("gap", C.c_double), ("energy_err_est", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# task-table element layouts (must match tjsq_plan.hpp)
SEG_DTYPE = np.dtype([("l_off", "<i8"), ("r_off", "<i8"), ("l_ld", "<i4"), ("r_ld", "<i4"), ("K", "<i4"),
                      ("l_space", "u1"), ("r_space", "u1"), ("l_t", "u1"), ("r_t", "u1"), ("c", "<f8")])
TILE_DTYPE = np.dtype([("out_off", "<i8"), ("out_ld", "<i4"), ("m0", "<i4"), ("n0", "<i4"), ("mt", "<i4"),
                       ("nt", "<i4"), ("seg_begin", "<i4"), ("seg_end", "<i4"), ("out_space", "<i4"), ("slot", "<i4"),
                       ("group", "<i4")])
MERGE_DTYPE = np.dtype([("dst_off", "<i8"), ("src_off", "<i8"), ("n", "<i8"), ("coef", "<f8")])
assert SEG_DTYPE.itemsize == 40 and TILE_DTYPE.itemsize == 48 and MERGE_DTYPE.itemsize == 32
# packed BF16x3 tables (tjsq_hostplan_packed)
PACKJOB_DTYPE = np.dtype([("src_off", "<i8"), ("dst_off", "<i8"), ("src_ld", "<i4"), ("rows", "<i4"),
                          ("cols", "<i4"), ("trans", "<i4"), ("lines", "<i4"), ("kpad", "<i4")])
PSEG_DTYPE = np.dtype([("l_off", "<i8"), ("r_off", "<i8"), ("l_lim", "<u4"), ("r_lim", "<u4"),
                       ("l_stride", "<i4"), ("r_stride", "<i4"), ("K", "<i4"), ("spaces", "<i4"), ("c", "<f8")])
PTILE_DTYPE = np.dtype([("out_off", "<i8"), ("out_ld", "<i4"), ("m0", "<i4"), ("n0", "<i4"), ("mb", "<i4"),
                        ("nb", "<i4"), ("mlim", "<i4"), ("nlim", "<i4"), ("seg_begin", "<i4"), ("seg_end", "<i4"),
                        ("out_kind", "<i4"), ("slot", "<i4"), ("pad0", "<i4")])
PREDUCE_DTYPE = np.dtype([("out_off", "<i8"), ("out_ld", "<i4"), ("m0", "<i4"), ("n0", "<i4"), ("mt", "<i4"),
                          ("nt", "<i4"), ("s0", "<i4"), ("s1", "<i4"), ("pad0", "<i4")])
assert (PACKJOB_DTYPE.itemsize, PSEG_DTYPE.itemsize, PTILE_DTYPE.itemsize, PREDUCE_DTYPE.itemsize) == (40, 48, 56, 40)

EXPORTS = ["tjsq_w6js", "tjsq_get_basis", "tjsq_get_super_basis", "tjsq_hmerge_tables", "tjsq_active_sites",
           "tjsq_plan_create", "tjsq_plan_stats_get", "tjsq_plan_destroy", "tjsq_heff_apply", "tjsq_heff_apply_f32",
           "tjsq_lanczos_solve", "tjsq_hostplan_create", "tjsq_hostplan_sizes", "tjsq_hostplan_tables",
           "tjsq_hostplan_stats", "tjsq_hostplan_destroy", "tjsq_plan_profile", "tjsq_plan_profile_read",
           "tjsq_heff_apply_host", "tjsq_lanczos_solve_host",
           "tjsq_block_update_create", "tjsq_trun_project_create", "tjsq_density_create", "tjsq_wave_transform_create",
           "tjsq_dia_transport_create", "tjsq_dia_transport_pairs_create", "tjsq_measure_create",
           "tjsq_w3js", "tjsq_ndia_transport_create", "tjsq_pair_ndia_create", "tjsq_cor_ndia_create",
           "tjsq_product_ndia_create", "tjsq_site_cor_create",
           "tjsq_program_sizes",
           "tjsq_program_layout", "tjsq_program_tables", "tjsq_program_stats", "tjsq_program_run",
           "tjsq_program_run_host", "tjsq_truncation_operator_host",
           "tjsq_program_profile", "tjsq_program_profile_read", "tjsq_program_destroy", "tjsq_truncation_operator",
           "tjsq_truncation_select", "tjsq_identity_truncation", "tjsq_batch_create", "tjsq_batch_offsets",
           "tjsq_batch_stats", "tjsq_batch_graph", "tjsq_batch_apply", "tjsq_batch_apply_f32", "tjsq_batch_profile",
           "tjsq_batch_profile_read", "tjsq_batch_destroy", "tjsq_plan_create_ex", "tjsq_hostplan_create_ex",
           "tjsq_plan_reshape", "tjsq_plan_graphs", "tjsq_shutdown", "tjsq_hostplan_packed", "tjsq_heff_apply_noalloc",
           "tjsq_plan_set_ops", "tjsq_last_error"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run grand_sql_tj_amd.build.build() (hipcc, gfx950) first")
        L = C.CDLL(LIB_PATH)
        L.tjsq_w6js.restype = C.c_double
        L.tjsq_w6js.argtypes = [C.c_int] * 6
        L.tjsq_w3js.restype = C.c_double
        L.tjsq_w3js.argtypes = [C.c_int] * 6
        L.tjsq_last_error.restype = C.c_char_p
        L.tjsq_last_error.argtypes = []
        P, I, I64 = C.c_void_p, C.c_int, C.c_int64
        sig = {
            "tjsq_get_basis": [I, I, P, I, P, P, P, P, P],
            "tjsq_get_super_basis": [P, P, I, I, P, P],
            "tjsq_hmerge_tables": [P, I, I, P, P, P],
            "tjsq_active_sites": [P, I],
            "tjsq_plan_create": [P, I, I, P, P, P, P, I, P, I, P, P],
            "tjsq_plan_stats_get": [P, P],
            "tjsq_plan_destroy": [P],
            "tjsq_heff_apply": [P, P, P, P],
            "tjsq_heff_apply_f32": [P, I, P, P, P],
            "tjsq_lanczos_solve": [P, P, P, P, P],
            "tjsq_hostplan_create": [P, I, I, P, P, P, P, I, P, P],
            "tjsq_hostplan_sizes": [P, P],
            "tjsq_hostplan_tables": [P, P, P, P, P, P, P],
            "tjsq_hostplan_stats": [P, P],
            "tjsq_hostplan_destroy": [P],
            "tjsq_plan_profile": [P, I],
            "tjsq_plan_profile_read": [P, P, P],
            "tjsq_heff_apply_host": [P, P, P],
            "tjsq_lanczos_solve_host": [P, P, P, P],
            "tjsq_block_update_create": [P, I, I, I, P, P, P, P],
            "tjsq_trun_project_create": [P, P, P, P],
            "tjsq_dia_transport_create": [P, P, P, P, P],
            "tjsq_dia_transport_pairs_create": [P, I, P, P, P, P],
            "tjsq_measure_create": [I, P, P, P, I, P, P],
            "tjsq_ndia_transport_create": [I, I, P, P, P, P, P],
            "tjsq_pair_ndia_create": [I, I, P, P, P, P, P],
            "tjsq_cor_ndia_create": [I, P, P, I, P, I, P, I, P, P, I, P, P],
            "tjsq_product_ndia_create": [I, I, I, P, P, P, P, P],
            "tjsq_site_cor_create": [I, I, P, I, P, I, P, I, P, P, I, P, P],
            "tjsq_density_create": [I, I, I, P, I, P],
            "tjsq_wave_transform_create": [I, P, P, P, P, I, P, I, P, P],
            "tjsq_program_sizes": [P, P],
            "tjsq_program_layout": [P, P, P, P],
            "tjsq_program_tables": [P, P, P, P, P, P],
            "tjsq_program_stats": [P, P],
            "tjsq_program_run": [P, P, P, P, P],
            "tjsq_program_profile": [P, I],
            "tjsq_program_profile_read": [P, P, P],
            "tjsq_program_destroy": [P],
            "tjsq_truncation_operator": [I, P, P, P, I, I, C.c_double, P, P, P, P, P, P],
            "tjsq_truncation_operator_host": [I, P, P, P, I, I, C.c_double, P, P, P, P, P],
            "tjsq_program_run_host": [P, P, P, P],
            "tjsq_truncation_select": [I, P, P, I, I, C.c_double, P, P, P, P],
            "tjsq_identity_truncation": [P, P, P, P, P, P],
            "tjsq_batch_create": [P, I, I, P, P],
            "tjsq_batch_offsets": [P, P],
            "tjsq_batch_stats": [P, P],
            "tjsq_batch_graph": [P, I],
            "tjsq_batch_apply": [P, P, P, P],
            "tjsq_batch_apply_f32": [P, I, P, P, P],
            "tjsq_batch_profile": [P, I],
            "tjsq_batch_profile_read": [P, P, P],
            "tjsq_batch_destroy": [P],
            "tjsq_plan_create_ex": [P, I, I, P, P, P, P, I, P, I, P, P, P],
            "tjsq_hostplan_create_ex": [P, I, I, P, P, P, P, I, P, P, P],
            "tjsq_plan_reshape": [P, P, P],
            "tjsq_plan_set_ops": [P, P, P, P],
            "tjsq_plan_graphs": [P, I],
            "tjsq_shutdown": [],
            "tjsq_heff_apply_noalloc": [P, I, P, P, P],
            "tjsq_hostplan_packed": [P, P, P, P, P, P, P, P, P],
        }
        for name in EXPORTS:
            f = getattr(L, name)
            if name in sig:
                f.restype = C.c_int
                f.argtypes = sig[name]
        del I64
        _lib = L
        # release the library's device state while the GPU runtime is still
        # up: atexit handlers run last-registered first, so this one runs
        # before those of torch and the HIP runtime registered earlier
        atexit.register(_shutdown)
    return _lib


_shut = False


def _shutdown():
    """tjsq_shutdown at interpreter exit; afterwards the finalizers of
    plans, programs and batches leave their handles alone (the process is
    ending, and the GPU runtime may already be gone when they run)."""
    global _shut
    if _lib is not None and not _shut:
        _shut = True
        _lib.tjsq_shutdown()


def _alive() -> bool:
    return _lib is not None and not _shut


def check(rc: int, what: str):
    if rc != TJSQ_OK:
        msg = lib().tjsq_last_error().decode()
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def lattice_struct(model) -> Lattice:
    lat = Lattice()
    lat.ny = model.ny
    lat.nx_tables = model.table_columns
    lat.jt[0], lat.jt[1] = model.jt
    lat.jd[0], lat.jd[1] = model.jd
    lat.jn[0], lat.jn[1] = model.jn
    lat.esite = model.esite
    lat.up_bias = model.up_bias
    return lat


class _Keep:
    """Holds the numpy arrays a C struct points to."""

    def __init__(self):
        self.arrays = []

    def ptr(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.arrays.append(a)
        return a.ctypes.data


def opset_struct(ops, keep: _Keep, device_data=None) -> OpSetC:
    s = OpSetC()
    s.length = ops.length
    s.nops = ops.nops
    s.hdr = keep.ptr(ops.hdr, np.int32)
    s.sub = keep.ptr(ops.sub, np.int32)
    s.off = keep.ptr(ops.off, np.int64)
    if device_data is not None:
        s.data = C.c_void_p(device_data)
        s.data_on_device = 1
    else:
        s.data = keep.ptr(ops.data, np.float64)
        s.data_on_device = 0
    s.ndata = int(ops.data.shape[0])
    return s


def basis_struct(bs, keep: _Keep) -> BasisC:
    b = BasisC()
    b.length = bs.length
    b.nsec = bs.nsec
    b.sec = keep.ptr(bs.sec, np.int32)
    b.sub = keep.ptr(bs.sub, np.int32)
    return b


def w6js(*j) -> float:
    return lib().tjsq_w6js(*[int(v) for v in j])


def w3js(*j) -> float:
    return lib().tjsq_w3js(*[int(v) for v in j])


def get_basis(block_len, blk, site):
    blk = np.ascontiguousarray(blk, dtype=np.int32)
    site = np.ascontiguousarray(site, dtype=np.int32)
    cap = 3 * max(1, blk.shape[0]) * max(1, site.shape[0])
    sec = np.zeros(4 * cap, np.int32)
    sub = np.zeros(6 * cap, np.int32)
    ns, nq = C.c_int(0), C.c_int(0)
    check(lib().tjsq_get_basis(block_len, blk.shape[0], blk.ctypes.data, site.shape[0], site.ctypes.data,
                               sec.ctypes.data, sub.ctypes.data, C.byref(ns), C.byref(nq)), "tjsq_get_basis")
    return sec[:4 * ns.value].reshape(-1, 4), sub[:6 * nq.value].reshape(-1, 6)


def get_super_basis(sys_bs, env_bs, total_up, total_down):
    keep = _Keep()
    s, e = basis_struct(sys_bs, keep), basis_struct(env_bs, keep)
    w = np.zeros(6 * max(1, sys_bs.nsec * env_bs.nsec), np.int32)
    nw = C.c_int(0)
    check(lib().tjsq_get_super_basis(C.byref(s), C.byref(e), total_up, total_down, w.ctypes.data, C.byref(nw)),
          "tjsq_get_super_basis")
    return w[:6 * nw.value].reshape(-1, 6)


def hmerge_tables(model, sys_len, env_len):
    lat = lattice_struct(model)
    cbs = np.zeros(4, np.int32)
    cbb = np.zeros(2, np.int32)
    inb = np.zeros(sys_len + env_len + 3, np.int32)
    check(lib().tjsq_hmerge_tables(C.byref(lat), sys_len, env_len, cbs.ctypes.data, cbb.ctypes.data,
                                   inb.ctypes.data), "tjsq_hmerge_tables")
    return cbs, cbb, inb


def active_sites(model, length) -> int:
    lat = lattice_struct(model)
    r = lib().tjsq_active_sites(C.byref(lat), length)
    if r < 0:
        check(r, "tjsq_active_sites")
    return r


class Shape(C.Structure):
    """Work-item shape of the grouped GEMM (tjsq_shape)."""
    _fields_ = [("big", C.c_int), ("small", C.c_int), ("k_split", C.c_int), ("item_chunks", C.c_int),
                ("xcds", C.c_int), ("merge_right", C.c_int), ("order", C.c_int)]
    DEFAULT = {"big": 64, "small": 32, "k_split": 256, "item_chunks": 32, "xcds": 1, "merge_right": 1, "order": 0}

    @classmethod
    def of(cls, d):
        if d is None:
            return None
        v = dict(cls.DEFAULT)
        v.update(d)
        return cls(**v)


def _shape_ptr(shape):
    sh = Shape.of(shape)
    return (C.byref(sh) if sh is not None else None), sh


class HostPlan:
    """Host-only compile of a superblock's H_eff into task tables."""

    def __init__(self, model, sb, shape=None):
        keep = _Keep()
        lat = lattice_struct(model)
        s, e = opset_struct(sb.sys, keep), opset_struct(sb.env, keep)
        sbs, ebs = basis_struct(sb.sys_bs, keep), basis_struct(sb.env_bs, keep)
        w = keep.ptr(sb.wsec, np.int32)
        h = C.c_void_p()
        shp, _sh = _shape_ptr(shape)
        check(lib().tjsq_hostplan_create_ex(C.byref(lat), sb.sys_len, sb.env_len, C.byref(s), C.byref(e),
                                            C.byref(sbs), C.byref(ebs), int(sb.wsec.shape[0]), C.c_void_p(w), shp,
                                            C.byref(h)), "tjsq_hostplan_create_ex")
        self.h = h
        sizes = np.zeros(15, np.int64)
        check(lib().tjsq_hostplan_sizes(h, sizes.ctypes.data), "tjsq_hostplan_sizes")
        (n1, t1, n2, t2, nm, self.arena_elems, self.scratch_elems, self.dim, self.sys_base, self.env_base,
         self.merged_base, nr, self.nslots, self.n64_1, self.n64_2) = [int(v) for v in sizes]
        self.seg1 = np.zeros(n1, SEG_DTYPE)
        self.tile1 = np.zeros(t1, TILE_DTYPE)
        self.seg2 = np.zeros(n2, SEG_DTYPE)
        self.tile2 = np.zeros(t2, TILE_DTYPE)
        self.merge = np.zeros(nm, MERGE_DTYPE)
        self.rtile = np.zeros(nr, TILE_DTYPE)
        check(lib().tjsq_hostplan_tables(h, self.seg1.ctypes.data, self.tile1.ctypes.data, self.seg2.ctypes.data,
                                         self.tile2.ctypes.data, self.merge.ctypes.data, self.rtile.ctypes.data),
              "tjsq_hostplan_tables")
        st = PlanStats()
        check(lib().tjsq_hostplan_stats(h, C.byref(st)), "tjsq_hostplan_stats")
        self.stats = st.as_dict()

    def packed(self):
        """The packed BF16x3 tables (dict of numpy arrays and byte sizes), or
        None where that layout does not apply."""
        sizes = np.zeros(11, np.int64)
        if lib().tjsq_hostplan_packed(self.h, sizes.ctypes.data, *([None] * 7)) != TJSQ_OK:
            return None
        n = [int(v) for v in sizes]
        out = {"ops_jobs": np.zeros(n[0], PACKJOB_DTYPE), "x_jobs": np.zeros(n[1], PACKJOB_DTYPE),
               "seg1": np.zeros(n[2], PSEG_DTYPE), "tile1": np.zeros(n[3], PTILE_DTYPE),
               "seg2": np.zeros(n[4], PSEG_DTYPE), "tile2": np.zeros(n[5], PTILE_DTYPE),
               "rtile": np.zeros(n[6], PREDUCE_DTYPE)}
        check(lib().tjsq_hostplan_packed(self.h, sizes.ctypes.data,
                                         *[out[k].ctypes.data for k in ("ops_jobs", "x_jobs", "seg1", "tile1", "seg2",
                                                                        "tile2", "rtile")]), "tjsq_hostplan_packed")
        out.update(nslots=n[7], ops_bytes=n[8], x_bytes=n[9], t_bytes=n[10])
        return out

    def __del__(self):
        if getattr(self, "h", None) and _alive and _alive():
            _lib.tjsq_hostplan_destroy(self.h)
            self.h = None


class Plan:
    """Device plan of one superblock (tjsq_plan_create).  Vectors are torch
    CUDA tensors; the stream is torch's current stream."""

    def __init__(self, model, sb, precisions=("fp64", "fp32", "bf16x3"), device_ops=None, shape=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("tjsq Plan needs a GPU (HIP device)")
        self.torch = torch
        keep = _Keep()
        lat = lattice_struct(model)
        dsys = device_ops[0].data_ptr() if device_ops is not None else None
        denv = device_ops[1].data_ptr() if device_ops is not None else None
        s, e = opset_struct(sb.sys, keep, dsys), opset_struct(sb.env, keep, denv)
        sbs, ebs = basis_struct(sb.sys_bs, keep), basis_struct(sb.env_bs, keep)
        w = keep.ptr(sb.wsec, np.int32)
        mask = 0
        for p in precisions:
            mask |= 1 << PRECISIONS[p]
        h = C.c_void_p()
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        shp, _sh = _shape_ptr(shape)
        check(lib().tjsq_plan_create_ex(C.byref(lat), sb.sys_len, sb.env_len, C.byref(s), C.byref(e), C.byref(sbs),
                                        C.byref(ebs), int(sb.wsec.shape[0]), C.c_void_p(w), mask, shp, stream,
                                        C.byref(h)), "tjsq_plan_create_ex")
        self.shape = shape
        self.h = h
        self.dim = sb.dim
        st = PlanStats()
        check(lib().tjsq_plan_stats_get(h, C.byref(st)), "tjsq_plan_stats_get")
        self.stats = st.as_dict()

    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def set_ops(self, sys_ops, env_ops, device_ops=None):
        """New block-operator data with the plan's layouts (tjsq_plan_set_ops):
        sys_ops / env_ops are OpSets whose hdr, sub and off equal those the
        plan was made with; their data (or device_ops, two CUDA tensors) are
        uploaded, hmerge and the FP32 / BF16x3 copies are redone."""
        keep = _Keep()
        dsys = device_ops[0].data_ptr() if device_ops is not None else None
        denv = device_ops[1].data_ptr() if device_ops is not None else None
        s, e = opset_struct(sys_ops, keep, dsys), opset_struct(env_ops, keep, denv)
        check(lib().tjsq_plan_set_ops(self.h, C.byref(s), C.byref(e), self._stream()), "tjsq_plan_set_ops")

    def graphs(self, enable=True):
        """Replay each Lanczos iteration from a captured hipGraph."""
        check(lib().tjsq_plan_graphs(self.h, int(enable)), "tjsq_plan_graphs")

    def reshape(self, shape):
        """Recompile the task tables for another work-item shape."""
        shp, _sh = _shape_ptr(shape)
        check(lib().tjsq_plan_reshape(self.h, shp, self._stream()), "tjsq_plan_reshape")
        st = PlanStats()
        check(lib().tjsq_plan_stats_get(self.h, C.byref(st)), "tjsq_plan_stats_get")
        self.stats = st.as_dict()
        self.shape = shape

    def apply(self, x, y=None, precision="fp64"):
        """y = H_eff x (model_wave) for a CUDA tensor x of length D."""
        torch = self.torch
        p = PRECISIONS[precision]
        dt = torch.float64 if p == FP64 else torch.float32
        if x.dtype != dt or not x.is_cuda or x.numel() != self.dim or not x.is_contiguous():
            raise ValueError(f"x must be a contiguous CUDA {dt} tensor of length {self.dim}")
        if y is None:
            y = torch.empty_like(x)
        if p == FP64:
            check(lib().tjsq_heff_apply(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), self._stream()),
                  "tjsq_heff_apply")
        else:
            check(lib().tjsq_heff_apply_f32(self.h, p, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                            self._stream()), "tjsq_heff_apply_f32")
        return y

    def lanczos(self, psi, tol=1e-10, max_iter=40, precision="fp64", refine_tol=1e-6, max_refine=1, energy_tol=0.0,
                rtol=0.0):
        """Ground state (lanczos_memo_new, levels = 1).  psi: CUDA float64
        start vector, overwritten with the normalised ground state."""
        torch = self.torch
        if psi.dtype != torch.float64 or not psi.is_cuda or psi.numel() != self.dim:
            raise ValueError("psi must be a CUDA float64 tensor of length D")
        o = LanczosOpts()
        o.tol, o.max_iter, o.precision = tol, max_iter, PRECISIONS[precision]
        o.refine_tol, o.max_refine = refine_tol, max_refine
        o.energy_tol, o.rtol = energy_tol, rtol
        r = LanczosResult()
        check(lib().tjsq_lanczos_solve(self.h, C.c_void_p(psi.data_ptr()), C.byref(o), C.byref(r), self._stream()),
              "tjsq_lanczos_solve")
        return r.as_dict()

    def profile(self, on=True):
        check(lib().tjsq_plan_profile(self.h, 1 if on else 0), "tjsq_plan_profile")

    def profile_read(self):
        """(ms, launches) per kernel kind since the last read: stage-1 64x64,
        stage-1 32x32, stage-2 64x64, stage-2 32x32, reduction."""
        ms = np.zeros(PROFILE_KINDS)
        n = np.zeros(PROFILE_KINDS, np.int64)
        check(lib().tjsq_plan_profile_read(self.h, ms.ctypes.data, n.ctypes.data), "tjsq_plan_profile_read")
        return ms, n

    def __del__(self):
        if getattr(self, "h", None) and _alive and _alive():
            _lib.tjsq_plan_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------------------
# density matrix, truncation and block update (tjsq_program)
# ---------------------------------------------------------------------------
# kernel kinds of profile_read: stage-1 64x64 items, stage-1 32x32, stage-2
# 64x64, stage-2 32x32, partial-tile reduction
PROFILE_KINDS = 5

PROGRAM_SIZE_KEYS = ("nops", "nsub", "ndata", "x_elems", "ops_elems", "n_seg1", "n_tile1", "n_seg2", "n_tile2",
                     "n_rtile", "n_slots", "scratch_elems", "n64_1", "n64_2")


class Program:
    """A compiled density-matrix, projection or block-update program.

    Built on the host (no GPU); ``run`` executes it on device pointers.  The
    output is an operator set whose layout ``layout()`` returns.
    """

    def __init__(self, handle):
        self._h = handle
        sizes = np.zeros(14, np.int64)
        check(lib().tjsq_program_sizes(self._h, sizes.ctypes.data), "tjsq_program_sizes")
        self.sizes = {k: int(v) for k, v in zip(PROGRAM_SIZE_KEYS, sizes)}

    @classmethod
    def block_update(cls, model, side, sys_len, env_len, pre, basis, trun):
        keep = _Keep()
        lat = lattice_struct(model)
        ps, bs, ts = opset_struct(pre, keep), basis_struct(basis, keep), opset_struct(trun, keep)
        h = C.c_void_p()
        check(lib().tjsq_block_update_create(C.byref(lat), side, sys_len, env_len, C.byref(ps), C.byref(bs),
                                             C.byref(ts), C.byref(h)), "tjsq_block_update_create")
        return cls(h)

    @classmethod
    def projection(cls, ops, ndia, trun):
        keep = _Keep()
        os_, ts = opset_struct(ops, keep), opset_struct(trun, keep)
        nd = np.ascontiguousarray(ndia, np.int32)
        h = C.c_void_p()
        check(lib().tjsq_trun_project_create(C.byref(os_), nd.ctypes.data, C.byref(ts), C.byref(h)),
              "tjsq_trun_project_create")
        return cls(h)

    @classmethod
    def dia_transport(cls, ops, site, basis, trun):
        """One step of the transport of diagonal operators along a block:
        ops (diagonal, block length L, or None) embedded into basis (length
        L + 1), the site operator (host data, or None) embedded after them,
        each projected with trun.  Run with ops = trun data, x = ops data."""
        keep = _Keep()
        os_ = opset_struct(ops, keep) if ops is not None and ops.nops > 0 else None
        ss = opset_struct(site, keep) if site is not None else None
        bs, ts = basis_struct(basis, keep), opset_struct(trun, keep)
        h = C.c_void_p()
        check(lib().tjsq_dia_transport_create(C.byref(os_) if os_ is not None else None,
                                              C.byref(ss) if ss is not None else None, C.byref(bs), C.byref(ts),
                                              C.byref(h)), "tjsq_dia_transport_create")
        return cls(h)

    @classmethod
    def dia_transport_pairs(cls, ops, nprod, site, basis, trun):
        """dia_transport with the products of the first nprod operators and
        the site operator appended; output order (nprod > 0): ops[:nprod],
        the site operator, ops[nprod:], the products."""
        keep = _Keep()
        os_ = opset_struct(ops, keep) if ops is not None and ops.nops > 0 else None
        ss = opset_struct(site, keep)
        bs, ts = basis_struct(basis, keep), opset_struct(trun, keep)
        h = C.c_void_p()
        check(lib().tjsq_dia_transport_pairs_create(C.byref(os_) if os_ is not None else None, int(nprod),
                                                    C.byref(ss), C.byref(bs), C.byref(ts), C.byref(h)),
              "tjsq_dia_transport_pairs_create")
        return cls(h)

    @classmethod
    def measure(cls, side, ops, site, basis, wsec):
        """<psi|O|psi> of diagonal block operators (or None) and the site
        operator of one side: the output is nops + 1 wavefunction-sized
        vectors y_j with <O_j> = psi . y_j.  Run with ops = the operators'
        data, x = psi, y zeroed."""
        keep = _Keep()
        os_ = opset_struct(ops, keep) if ops is not None and ops.nops > 0 else None
        ss = opset_struct(site, keep) if site is not None else None
        bs = basis_struct(basis, keep)
        w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
        h = C.c_void_p()
        check(lib().tjsq_measure_create(side, C.byref(os_) if os_ is not None else None,
                                        C.byref(ss) if ss is not None else None, C.byref(bs), w.shape[0],
                                        w.ctypes.data, C.byref(h)), "tjsq_measure_create")
        return cls(h)

    @classmethod
    def ndia_transport(cls, side, flag, ops, site, basis, trun):
        """One step of the transport of rank-k operators along a block (side
        0 system, 1 environment; flag 'B' or 'F'): ops (block length L, or
        None) embedded into basis (length L + 1), the site operator (host
        data, or None) embedded after them, each projected with trun.  Run
        with ops = trun data, x = ops data."""
        keep = _Keep()
        os_ = opset_struct(ops, keep) if ops is not None and ops.nops > 0 else None
        ss = opset_struct(site, keep) if site is not None else None
        bs, ts = basis_struct(basis, keep), opset_struct(trun, keep)
        h = C.c_void_p()
        check(lib().tjsq_ndia_transport_create(int(side), ord(flag), C.byref(os_) if os_ is not None else None,
                                               C.byref(ss) if ss is not None else None, C.byref(bs), C.byref(ts),
                                               C.byref(h)), "tjsq_ndia_transport_create")
        return cls(h)

    @classmethod
    def pair_ndia(cls, side, flag, ops, site, basis, trun):
        """The rank-0 products [O x s]^0 of every rank-k block operator of
        ops (length L - 1) with the site operator joining at length L,
        projected with trun: diagonal operators at length L.  Run with
        ops = trun data, x = ops data."""
        keep = _Keep()
        os_, ss = opset_struct(ops, keep), opset_struct(site, keep)
        bs, ts = basis_struct(basis, keep), opset_struct(trun, keep)
        h = C.c_void_p()
        check(lib().tjsq_pair_ndia_create(int(side), ord(flag), C.byref(os_), C.byref(ss), C.byref(bs),
                                          C.byref(ts), C.byref(h)), "tjsq_pair_ndia_create")
        return cls(h)

    @classmethod
    def cor_ndia(cls, flag, sys_op, env_ops, env_trans, env_site, env_site_trans, sys_site, sys_site_trans,
                 sys_basis, env_basis, wsec):
        """Correlations across the bond of one system block operator with
        environment block operators, the environment bond-site operator and
        the system bond-site operator (each optional): one
        wavefunction-sized output vector y_q per partner, value = psi . y_q.
        Run with ops = concat(sys_op data, env_ops data), x = psi, y zeroed."""
        keep = _Keep()
        so = opset_struct(sys_op, keep)
        eo = opset_struct(env_ops, keep) if env_ops is not None and env_ops.nops > 0 else None
        es = opset_struct(env_site, keep) if env_site is not None else None
        ss = opset_struct(sys_site, keep) if sys_site is not None else None
        sb, eb = basis_struct(sys_basis, keep), basis_struct(env_basis, keep)
        w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
        ref = lambda v: C.byref(v) if v is not None else None
        h = C.c_void_p()
        check(lib().tjsq_cor_ndia_create(ord(flag), C.byref(so), ref(eo), int(env_trans), ref(es),
                                         int(env_site_trans), ref(ss), int(sys_site_trans), C.byref(sb),
                                         C.byref(eb), w.shape[0], w.ctypes.data, C.byref(h)),
              "tjsq_cor_ndia_create")
        return cls(h)

    @classmethod
    def product_ndia(cls, side, flag, rank, ops, site, basis, trun):
        """The couplings [O x s]^rank of every operator of ops (length L - 1)
        with the site operator joining at length L, projected with trun:
        operators at length L.  Run with ops = trun data, x = ops data."""
        keep = _Keep()
        os_, ss = opset_struct(ops, keep), opset_struct(site, keep)
        bs, ts = basis_struct(basis, keep), opset_struct(trun, keep)
        h = C.c_void_p()
        check(lib().tjsq_product_ndia_create(int(side), ord(flag), int(rank), C.byref(os_), C.byref(ss),
                                             C.byref(bs), C.byref(ts), C.byref(h)), "tjsq_product_ndia_create")
        return cls(h)

    @classmethod
    def site_cor(cls, kind, flag, sys_site, sys_site_trans, env_ops, env_trans, env_site, env_site_trans, sys_basis,
                 env_basis, wsec):
        """The bond-site contractions of Get_SC_Order: kind 0 system site x
        environment site, 1 system site x each environment block operator,
        2 each environment block operator x environment site.  Run with
        ops = env_ops data (kind 0: any), x = psi, y zeroed."""
        keep = _Keep()
        ss = opset_struct(sys_site, keep) if sys_site is not None else None
        eo = opset_struct(env_ops, keep) if env_ops is not None and env_ops.nops > 0 else None
        es = opset_struct(env_site, keep) if env_site is not None else None
        sb, eb = basis_struct(sys_basis, keep), basis_struct(env_basis, keep)
        w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
        ref = lambda v: C.byref(v) if v is not None else None
        h = C.c_void_p()
        check(lib().tjsq_site_cor_create(int(kind), ord(flag), ref(ss), int(sys_site_trans), ref(eo), int(env_trans),
                                         ref(es), int(env_site_trans), C.byref(sb), C.byref(eb), w.shape[0],
                                         w.ctypes.data, C.byref(h)), "tjsq_site_cor_create")
        return cls(h)

    @classmethod
    def wave_transform(cls, direction, sys_bs, env_bs, sys_trun, env_trun, wsec_old, wsec_new):
        """wave_trans_LR (direction 0) / wave_trans_RL (1): run with ops =
        concat(sys_trun.data, env_trun.data), x = old psi, y = new psi."""
        keep = _Keep()
        sb, eb = basis_struct(sys_bs, keep), basis_struct(env_bs, keep)
        st, et = opset_struct(sys_trun, keep), opset_struct(env_trun, keep)
        wo = np.ascontiguousarray(wsec_old, np.int32).reshape(-1, 6)
        wn = np.ascontiguousarray(wsec_new, np.int32).reshape(-1, 6)
        h = C.c_void_p()
        check(lib().tjsq_wave_transform_create(direction, C.byref(sb), C.byref(eb), C.byref(st), C.byref(et),
                                               wo.shape[0], wo.ctypes.data, wn.shape[0], wn.ctypes.data,
                                               C.byref(h)), "tjsq_wave_transform_create")
        return cls(h)

    @classmethod
    def density(cls, side, length, wsec, levels=1):
        w = np.ascontiguousarray(wsec, np.int32).reshape(-1, 6)
        h = C.c_void_p()
        check(lib().tjsq_density_create(side, length, w.shape[0], w.ctypes.data, levels, C.byref(h)),
              "tjsq_density_create")
        return cls(h)

    def layout(self):
        """(hdr (nops, 4), sub (nsub, 5), off (nsub,)) of the output operator set."""
        hdr = np.zeros((self.sizes["nops"], 4), np.int32)
        sub = np.zeros((self.sizes["nsub"], 5), np.int32)
        off = np.zeros(self.sizes["nsub"], np.int64)
        check(lib().tjsq_program_layout(self._h, hdr.ctypes.data, sub.ctypes.data, off.ctypes.data),
              "tjsq_program_layout")
        return hdr, sub, off

    def tables(self):
        z = self.sizes
        seg1 = np.zeros(z["n_seg1"], SEG_DTYPE)
        tile1 = np.zeros(z["n_tile1"], TILE_DTYPE)
        seg2 = np.zeros(z["n_seg2"], SEG_DTYPE)
        tile2 = np.zeros(z["n_tile2"], TILE_DTYPE)
        rtile = np.zeros(z["n_rtile"], TILE_DTYPE)
        check(lib().tjsq_program_tables(self._h, seg1.ctypes.data, tile1.ctypes.data, seg2.ctypes.data,
                                        tile2.ctypes.data, rtile.ctypes.data), "tjsq_program_tables")
        return {"seg1": seg1, "tile1": tile1, "seg2": seg2, "tile2": tile2, "rtile": rtile}

    def stats(self):
        st = PlanStats()
        check(lib().tjsq_program_stats(self._h, C.byref(st)), "tjsq_program_stats")
        return st.as_dict()

    def run(self, ops, x, y, stream=None):
        """ops / x / y: torch CUDA float64 tensors (ops may be None for density programs)."""
        import torch
        assert y.dtype == torch.float64 and y.is_cuda and y.numel() >= self.sizes["ndata"]
        assert x.dtype == torch.float64 and x.numel() >= self.sizes["x_elems"]
        if ops is not None:
            assert ops.dtype == torch.float64 and ops.numel() >= self.sizes["ops_elems"]
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().tjsq_program_run(self._h, None if ops is None else ops.data_ptr(), x.data_ptr(), y.data_ptr(),
                                     C.c_void_p(s)), "tjsq_program_run")

    def profile(self, enable=True):
        check(lib().tjsq_program_profile(self._h, int(enable)), "tjsq_program_profile")

    def profile_read(self):
        ms = np.zeros(PROFILE_KINDS, np.float64)
        n = np.zeros(PROFILE_KINDS, np.int64)
        check(lib().tjsq_program_profile_read(self._h, ms.ctypes.data, n.ctypes.data), "tjsq_program_profile_read")
        return ms, n

    def opset(self, length, data):
        """blocks.OpSet over the output layout with host data ``data``."""
        from .blocks import OpSet
        hdr, sub, off = self.layout()
        return OpSet(length, hdr, sub, off, np.ascontiguousarray(data, np.float64))

    def __del__(self):
        # at interpreter exit the module's globals may already be gone
        if getattr(self, "_h", None) and _alive and _alive():
            _lib.tjsq_program_destroy(self._h)
            self._h = None


def truncation_select(den_sub, eigval, kept_min, kept_max, trun_err):
    """Host selection of Get_truncation_operator on given eigenvalues."""
    den_sub = np.ascontiguousarray(den_sub, np.int32).reshape(-1, 5)
    ev = np.ascontiguousarray(eigval, np.float64)
    n = den_sub.shape[0]
    counts = np.zeros(n, np.int32)
    spec = np.zeros(ev.shape[0], np.float64)
    order = np.zeros(ev.shape[0], np.int32)
    info = TruncInfo()
    check(lib().tjsq_truncation_select(n, den_sub.ctypes.data, ev.ctypes.data, kept_min, kept_max,
                                       C.c_double(trun_err), counts.ctypes.data, spec.ctypes.data, order.ctypes.data,
                                       C.byref(info)), "tjsq_truncation_select")
    return counts, spec, order, info.as_dict()


def truncation_operator(den_sub, den_off, den, kept_min, kept_max, trun_err, stream=None):
    """Get_truncation_operator on a device density matrix ``den`` (torch
    float64, overwritten by the eigenvectors).  Returns (trun_sub, trun_off,
    spectrum, info); the truncation operator's data is ``den``."""
    import torch
    den_sub = np.ascontiguousarray(den_sub, np.int32).reshape(-1, 5)
    den_off = np.ascontiguousarray(den_off, np.int64)
    n = den_sub.shape[0]
    tsub = np.zeros((n, 5), np.int32)
    toff = np.zeros(n, np.int64)
    nt = C.c_int(0)
    spec = np.zeros(int(den_sub[:, 4].sum()), np.float64)
    info = TruncInfo()
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().tjsq_truncation_operator(n, den_sub.ctypes.data, den_off.ctypes.data, den.data_ptr(), kept_min,
                                         kept_max, C.c_double(trun_err), tsub.ctypes.data, toff.ctypes.data,
                                         C.byref(nt), spec.ctypes.data, C.byref(info), C.c_void_p(s)),
          "tjsq_truncation_operator")
    return tsub[:nt.value].copy(), toff[:nt.value].copy(), spec, info.as_dict()


def identity_truncation(basis, data=None, stream=None):
    """Identity truncation layout over the basis sectors; fills ``data``
    (device float64) when given.  Returns (trun_sub, trun_off, ndata)."""
    keep = _Keep()
    b = basis_struct(basis, keep)
    tsub = np.zeros((basis.nsec, 5), np.int32)
    toff = np.zeros(basis.nsec, np.int64)
    nd = C.c_int64(0)
    ptr, s = None, None
    if data is not None:
        import torch
        ptr = data.data_ptr()
        s = C.c_void_p(stream if stream is not None else torch.cuda.current_stream().cuda_stream)
    check(lib().tjsq_identity_truncation(C.byref(b), tsub.ctypes.data, toff.ctypes.data, C.byref(nd), ptr, s),
          "tjsq_identity_truncation")
    return tsub, toff, nd.value


class Batch:
    """Applies of several plans in one grouped launch per stage (tjsq_batch).

    Vectors are the members' vectors back to back at ``offsets``."""

    def __init__(self, plans, graph=False, precisions=("fp64", "fp32", "bf16x3")):
        import torch
        hs = (C.c_void_p * len(plans))(*[p.h.value for p in plans])
        self._h = C.c_void_p()
        s = torch.cuda.current_stream().cuda_stream
        mask = 0
        for pr in precisions:
            mask |= 1 << PRECISIONS[pr]
        check(lib().tjsq_batch_create(hs, len(plans), mask, C.c_void_p(s), C.byref(self._h)), "tjsq_batch_create")
        off = np.zeros(len(plans) + 1, np.int64)
        check(lib().tjsq_batch_offsets(self._h, off.ctypes.data), "tjsq_batch_offsets")
        self.offsets = off
        self.dim = int(off[-1])
        self.graph(graph)

    def graph(self, enable=True):
        check(lib().tjsq_batch_graph(self._h, int(enable)), "tjsq_batch_graph")

    @property
    def stats(self):
        st = PlanStats()
        check(lib().tjsq_batch_stats(self._h, C.byref(st)), "tjsq_batch_stats")
        return st.as_dict()

    def apply(self, x, y=None, precision="fp64", stream=None):
        import torch
        if y is None:
            y = torch.empty_like(x)
        assert x.numel() == self.dim and y.numel() == self.dim and x.is_cuda and y.is_cuda
        s = C.c_void_p(stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        prec = PRECISIONS[precision]
        if prec == FP64:
            assert x.dtype == torch.float64 and y.dtype == torch.float64
            check(lib().tjsq_batch_apply(self._h, x.data_ptr(), y.data_ptr(), s), "tjsq_batch_apply")
        else:
            assert x.dtype == torch.float32 and y.dtype == torch.float32
            check(lib().tjsq_batch_apply_f32(self._h, prec, x.data_ptr(), y.data_ptr(), s), "tjsq_batch_apply_f32")
        return y

    def profile(self, enable=True):
        check(lib().tjsq_batch_profile(self._h, int(enable)), "tjsq_batch_profile")

    def profile_read(self):
        ms = np.zeros(PROFILE_KINDS, np.float64)
        n = np.zeros(PROFILE_KINDS, np.int64)
        check(lib().tjsq_batch_profile_read(self._h, ms.ctypes.data, n.ctypes.data), "tjsq_batch_profile_read")
        return ms, n

    def __del__(self):
        # at interpreter exit the module's globals may already be gone
        if getattr(self, "_h", None) and _alive and _alive():
            _lib.tjsq_batch_destroy(self._h)
            self._h = None
