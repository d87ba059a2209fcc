"""C-ABI library on the CPU: symbols, host tables and the compiled plan."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
import plan_eval
from helpers import GEOMS, ROOT, model, rel, small_superblock
from grand_sql_tj_amd import blocks, capi, synth


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tjsq_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tjsq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = capi.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.EXPORTS)


def test_w6js_bit_exact_with_oracle():
    rng = np.random.default_rng(3)
    for _ in range(3000):
        j = rng.integers(0, 11, size=6)
        assert capi.w6js(*j) == oracle.w6js(*j)


@pytest.mark.parametrize("ub", [0, 1])
def test_get_basis_and_super_basis_match_oracle(ub):
    rng = np.random.default_rng(11 + ub)
    site = blocks.site_sectors(ub)
    for _ in range(30):
        secs = set()
        for _ in range(rng.integers(1, 9)):
            s2 = int(rng.integers(0, 7))
            secs.add((s2 + 2 * int(rng.integers(0, 3)) if ub else 0, s2, int(rng.integers(1, 5))))
        blk = np.array(sorted(secs, key=lambda t: (-t[1], -t[0])), dtype=np.int32)
        cs, cq = capi.get_basis(4, blk, site)
        os_, oq = oracle.oracle_get_basis(4, blk, site)
        assert np.array_equal(cs, os_) and np.array_equal(cq, oq)
        b = blocks.Basis(5, cs, cq)
        for tot in (0, 2, 4):
            w = capi.get_super_basis(b, b, tot, 0)
            assert np.array_equal(w, blocks.get_super_basis(b, b, tot, 0))


@pytest.mark.parametrize("geom", GEOMS)
def test_hmerge_tables_match_oracle(geom):
    m = model()
    a = capi.hmerge_tables(m, *geom)
    b = oracle.hmerge_tables(m, *geom)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("ub", [1, 0])
@pytest.mark.parametrize("geom", GEOMS)
def test_compiled_plan_reproduces_oracle(ub, geom):
    # the task tables the device executes, evaluated with numpy
    m = model(up_bias=ub)
    sb = small_superblock(m, *geom, seed=7 * geom[0] + geom[1] + ub, msys=40, menv=44, median=5)
    x = synth.random_vector(sb, 9)
    hp = capi.HostPlan(m, sb)
    assert rel(plan_eval.evaluate(hp, sb, x), oracle.model_wave(m, sb, x)) < 1e-13
    st = hp.stats
    assert st["dim"] == sb.dim and st["flops_exec"] <= st["flops_ref"] * 1.0000001 + 1


def test_plan_rejects_inconsistent_operator_dims():
    m = model()
    sb = small_superblock(m, 7, 9, seed=2)
    bad = blocks.OpSet(sb.sys.length, sb.sys.hdr.copy(), sb.sys.sub.copy(), sb.sys.off.copy(), sb.sys.data)
    bad.sub[0, 3] += 1      # ham block row count no longer matches the basis
    bad.data = np.concatenate([bad.data, np.zeros(64)])
    bsb = blocks.Superblock(sb.sys_len, sb.env_len, bad, sb.env, sb.sys_bs, sb.env_bs, sb.wsec)
    with pytest.raises(RuntimeError, match="dimensions"):
        capi.HostPlan(m, bsb)


def test_plan_rejects_lattice_too_small():
    m = model()
    sb = small_superblock(m, 7, 9, seed=2)
    small = model(nx=2)
    with pytest.raises(RuntimeError, match="lattice too small"):
        capi.HostPlan(small, sb)


# every work-item shape the autotuner may choose compiles to tables that
# reproduce the oracle (blocks large enough for 64-tiles, K splits and
# partial tiles)
TUNE_SHAPES = [None, {"big": 32}, {"k_split": 32, "item_chunks": 1}, {"k_split": 64, "item_chunks": 4},
               {"k_split": 1024, "item_chunks": 64}, {"xcds": 8}, {"big": 32, "k_split": 128, "xcds": 4},
               {"merge_right": 1}, {"merge_right": 1, "big": 32, "k_split": 64, "item_chunks": 2, "xcds": 8},
               {"order": 1}, {"order": 1, "xcds": 8, "k_split": 64, "item_chunks": 2},
               {"order": 1, "merge_right": 0, "big": 32}]


@pytest.mark.parametrize("shape", TUNE_SHAPES)
def test_plan_shapes_reproduce_oracle(shape):
    m = model()
    sb = small_superblock(m, 11, 11, seed=23, msys=240, menv=220, median=30, nsec=8)
    x = synth.random_vector(sb, 4)
    hp = capi.HostPlan(m, sb, shape=shape)
    assert rel(plan_eval.evaluate(hp, sb, x), oracle.model_wave(m, sb, x)) < 1e-13
    if shape and shape.get("big") == 32:
        assert hp.n64_1 == 0 and hp.n64_2 == 0


def test_plan_rejects_unsupported_shape():
    m = model()
    sb = small_superblock(m, 7, 9, seed=2)
    for bad in ({"big": 128}, {"small": 16}, {"k_split": 8}, {"item_chunks": 0}, {"xcds": 0}, {"order": 2}):
        with pytest.raises(RuntimeError, match="shape"):
            capi.HostPlan(m, sb, shape=bad)


def test_serving_footprint_counts_the_plan():
    from grand_sql_tj_amd import serve
    m = model()
    small = small_superblock(m, 7, 9, seed=2, msys=40, menv=44, median=5)
    big = small_superblock(m, 7, 9, seed=2, msys=160, menv=150, median=14)
    hp = capi.HostPlan(m, big)
    f64 = serve.footprint(m, big, "fp64", 40)
    f32 = serve.footprint(m, big, "fp32", 40)
    assert f64 >= 8 * (hp.arena_elems + hp.scratch_elems) + 42 * big.dim * 8
    assert f32 > f64 - 42 * big.dim * 4          # extra FP32 operands outweigh the smaller basis
    assert serve.footprint(m, small, "fp64", 40) < f64


def test_host_buffer_entry_points_reject_bad_arguments():
    """The host-buffer forms the Fortran drop-in calls check their arguments
    before any device work (no GPU is touched here)."""
    lib = capi.lib()
    lib.tjsq_program_run_host.restype = C.c_int
    lib.tjsq_program_run_host.argtypes = [C.c_void_p] * 4
    assert lib.tjsq_program_run_host(None, None, None, None) != 0
    assert b"tjsq_program_run_host" in lib.tjsq_last_error()
    lib.tjsq_truncation_operator_host.restype = C.c_int
    sub = np.array([0, 0, 0, 2, 2], np.int32)
    off = np.zeros(1, np.int64)
    assert lib.tjsq_truncation_operator_host(0, sub.ctypes.data, off.ctypes.data, None, 1, 2, C.c_double(0.0),
                                             None, None, None, None, None) != 0
    assert b"tjsq_truncation_operator_host" in lib.tjsq_last_error()
