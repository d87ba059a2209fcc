"""Two-site SU(2) DMRG on the device: the warm-up and the finite sweeps of
the reference (warm_up_point, left_to_right, right_to_left;
code/tjsq/dmrg_sub.f90:556-1087, scheduled as Get_Ground_State,
code/tjsq/main.f90:96-187), assembled from the library's device calls.

Each step at a bond [sys | site | site | env]:

    tjsq_plan_create (hmerge)          -> tjsq_lanczos_solve (lanczos_memo_new)
    tjsq_density_create / _run         -> tjsq_truncation_operator (or identity)
    tjsq_block_update_create / _run    (Get_Hamiltonian_Sys / _Env)
    tjsq_wave_transform_create / _run  (wave_trans_LR / _RL, the next start vector)

Block models, truncation operators and the fused bases are kept per block
length, as the reference keeps them on disk (model_to_disk, truns_to_disk,
basis_to_disk); here the operator data stay in device memory and only their
layouts live on the host.  The reference's random start vectors
(wave_initialize, unseeded random_number) are seeded here.

Two schedules:
- run(sweeps): warm-up to half the lattice (the blocks grow together, with
  truncation once an enlarged block exceeds kept_max), then left-to-right
  and right-to-left half sweeps between the block lengths 1 and N - 3,
  `sweeps` times, every step with the tol, kept_max, kept_min and trun_err
  given;
- ground_state(...): Get_Ground_State's schedule (main.f90:96-187, a fresh
  run: restart1 = restart2 = 0, iter1 = 0) with its ramps of the Lanczos
  tolerance (lanerr x 1000 in the warm-up, x 30 in the first two half sweeps,
  then lanerr), of the truncation error (trun_err x 200, x 15, then
  trun_err) and of kept_max (kept_max, then 2/3 kept_max + 1/3 kept_maxf,
  then kept_maxf, with kept_min = min(3 kept_max / 4, 3600)), and its closing
  left-to-right pass to the block length Num_site0 / 2 + 3.
"""
from __future__ import annotations

import math
import os
import re
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .blocks import Basis, OpSet, OpSetBuilder, Superblock, get_basis, get_super_basis, site_sectors


def site_model(model) -> OpSet:
    """Block model of one site (Get_hamiltonian_site with the site operators
    of Get_site_operator_su2_U0, dmrg_sub.f90:370-517): the Hamiltonian is
    esite * n (warm_up_point, dmrg_sub.f90:576-580), one active site."""
    ub = model.up_bias
    s2 = math.sqrt(2.0)
    b = OpSetBuilder(1)
    b.add(0, 0, [(0, 0, 0, [[0.0]]), (ub, 1, 0, [[model.esite * s2]])])
    b.add(ub, 1, [(0, 0, 1, [[-s2]])])                 # elec_up: c+ on the empty site
    b.add(-ub, 1, [(ub, 1, -1, [[-s2]])])              # elec_down
    b.add(0, 2, [(ub, 1, 0, [[math.sqrt(1.5)]])])      # spin_sd
    b.add(0, 0, [(0, 0, 0, [[0.0]]), (ub, 1, 0, [[s2]])])   # num_sn
    return b.build()


@dataclass
class DeviceBlock:
    """A block model: operator-set layout on the host, data on the device."""
    ops: OpSet            # layout (its data array is a placeholder of the right size)
    data: object          # torch float64 CUDA tensor

    @property
    def length(self):
        return self.ops.length


@dataclass
class DeviceTrun:
    ops: OpSet            # one-operator layout
    data: object


@dataclass
class StepRecord:
    phase: str
    sys_len: int
    env_len: int
    dim: int
    energy: float
    iterations: int
    kept: int
    trunc_err: float
    seconds: float


@dataclass
class DMRG:
    model: object
    total_up: int
    kept_max: int
    kept_min: int = 0
    trun_err: float = 1e-8
    tol: float = 1e-10
    max_iter: int = 40
    precision: str = "fp64"
    seed: int = 1
    sys: dict = field(default_factory=dict)       # length -> DeviceBlock
    env: dict = field(default_factory=dict)
    sys_trun: dict = field(default_factory=dict)  # length L -> truncation that made block L
    env_trun: dict = field(default_factory=dict)
    sys_basis: dict = field(default_factory=dict)  # length L -> fused basis (block L-1 + site)
    env_basis: dict = field(default_factory=dict)
    log: list = field(default_factory=list)
    # seconds per phase over the run: plan (host compile + table upload),
    # solve (Lanczos), truncate (density matrix + eigensolve + selection),
    # grow (block update), transform (wave_trans_LR/RL); each phase ends
    # with a device synchronisation
    phase_s: dict = field(default_factory=lambda: {k: 0.0 for k in ("plan", "solve", "truncate", "grow", "transform")})

    def __post_init__(self):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("the device DMRG needs a GPU")
        self.torch = torch
        self.nsite = self.model.nsite
        self.site = site_sectors(self.model.up_bias)
        self.rng = np.random.default_rng(self.seed)
        self.psi = None          # ground state of the last bond (torch float64)
        self.wsec = None         # its sector list
        self.bond = None         # (sys_len, env_len) of that bond
        if self.kept_min <= 0:
            self.kept_min = min(3 * self.kept_max // 4, 3600)

    # ---- helpers ---------------------------------------------------------
    def _device(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, np.float64)).to("cuda")

    def _block(self, ops: OpSet) -> DeviceBlock:
        return DeviceBlock(OpSet(ops.length, ops.hdr, ops.sub, ops.off, np.zeros(ops.data.size)), self._device(ops.data))

    def _basis(self, blk: DeviceBlock) -> Basis:
        """Get_basis of the block with one site (the basis carries length + 1)."""
        return get_basis(blk.length, blk.ops.sectors(), self.site)

    def _timed(self, name, fn, *args):
        t0 = time.perf_counter()
        out = fn(*args)
        self.torch.cuda.synchronize()
        self.phase_s[name] += time.perf_counter() - t0
        return out

    def _solve(self, blk_s, blk_e, sys_bs, env_bs, wsec, psi0, phase):
        torch = self.torch
        sb = Superblock(blk_s.length, blk_e.length, blk_s.ops, blk_e.ops, sys_bs, env_bs, wsec, meta={})
        t0 = time.perf_counter()
        plan = self._timed("plan", lambda: capi.Plan(self.model, sb, precisions=("fp64", self.precision),
                                                     device_ops=(blk_s.data, blk_e.data)))
        if psi0 is None:
            psi = self._device(self.rng.standard_normal(sb.dim))
        else:
            psi = psi0.clone()
        psi /= psi.norm()
        info = self._timed("solve", lambda: plan.lanczos(psi, tol=self.tol, max_iter=self.max_iter,
                                                         precision=self.precision))
        self.psi, self.wsec = psi, wsec
        self.bond = (blk_s.length, blk_e.length)
        return info, time.perf_counter() - t0

    def _truncate(self, side, sys_len, env_len, wsec, basis, psi, allow_identity):
        """rho of one side, then its truncation operator (identity while the
        enlarged block fits in kept_max, as the warm-up does)."""
        torch = self.torch
        length = (sys_len if side == 0 else env_len) + 1
        if allow_identity and int(basis.sec[:, 2].sum()) <= self.kept_max:
            tsub, toff, nd = capi.identity_truncation(basis)
            data = torch.empty(max(1, nd), dtype=torch.float64, device="cuda")
            capi.identity_truncation(basis, data)
            kept, err = int(basis.sec[:, 2].sum()), 0.0
        else:
            den = capi.Program.density(side, length, wsec, 1)
            rho = torch.empty(max(1, den.sizes["ndata"]), dtype=torch.float64, device="cuda")
            den.run(None, psi, rho)
            _, dsub, doff = den.layout()
            tsub, toff, _, info = capi.truncation_operator(dsub, doff, rho, self.kept_min, self.kept_max,
                                                           self.trun_err)
            data, kept, err = rho, info["num_state"], info["trunc_err"]
        n = int((toff + tsub[:, 3].astype(np.int64) * tsub[:, 4]).max()) if tsub.shape[0] else 0
        ops = OpSet(length, np.array([[length, 0, 0, tsub.shape[0]]], np.int32), tsub.astype(np.int32),
                    toff.astype(np.int64), np.zeros(max(1, n)))
        return DeviceTrun(ops, data), kept, err

    def _grow(self, side, blk, basis, trun, sys_len, env_len) -> DeviceBlock:
        """Get_Hamiltonian_Sys (side 0) / _Env (side 1): the block one site longer."""
        torch = self.torch
        pg = capi.Program.block_update(self.model, side, sys_len, env_len, blk.ops, basis, trun.ops)
        hdr, sub, off = pg.layout()
        out = torch.empty(max(1, pg.sizes["ndata"]), dtype=torch.float64, device="cuda")
        pg.run(trun.data, blk.data, out)
        ops = OpSet(blk.length + 1, hdr, sub, off, np.zeros(max(1, pg.sizes["ndata"])))
        return DeviceBlock(ops, out)

    def _transform(self, direction, sys_bs, env_bs, sys_trun, env_trun, wsec_new):
        torch = self.torch
        pg = capi.Program.wave_transform(direction, sys_bs, env_bs, sys_trun.ops, env_trun.ops, self.wsec, wsec_new)
        ns = sys_trun.ops.data.size
        ops = torch.cat([sys_trun.data[:ns], env_trun.data[:env_trun.ops.data.size]])
        out = torch.empty(max(1, pg.sizes["ndata"]), dtype=torch.float64, device="cuda")
        pg.run(ops, self.psi, out)
        return out[:pg.sizes["ndata"]]

    def _record(self, phase, sl, el, wsec, info, kept, err, secs):
        d = int((wsec[:, 4].astype(np.int64) * wsec[:, 5]).sum())
        self.log.append(StepRecord(phase, sl, el, d, info["e0"], info["iterations"], kept, err, secs))

    # ---- warm_up_point (dmrg_sub.f90:556-725), bg_id = 1 ------------------
    def warm_up(self):
        site = site_model(self.model)
        self.sys[1] = self._block(site)
        self.env[1] = self._block(site)
        label, end = 1, self.nsite // 2
        while label + 1 <= end:
            s, e = self.sys[label], self.env[label]
            sys_bs, env_bs = self._basis(s), self._basis(e)
            tot = (sys_bs.length + env_bs.length)
            tot_up = int(1.0 * tot / self.nsite * self.total_up + 0.001)
            tot_up -= tot_up % 2
            tot_up = min(tot_up, self.total_up)
            wsec = get_super_basis(sys_bs, env_bs, tot_up, 0)
            info, secs = self._solve(s, e, sys_bs, env_bs, wsec, None, "warm-up")
            t_s, k_s, err = self._timed("truncate", self._truncate, 0, label, label, wsec, sys_bs, self.psi, True)
            t_e, _, _ = self._timed("truncate", self._truncate, 1, label, label, wsec, env_bs, self.psi, True)
            self._record("warm-up", label, label, wsec, info, k_s, err, secs)
            self.sys[label + 1] = self._timed("grow", self._grow, 0, s, sys_bs, t_s, label, label)
            self.env[label + 1] = self._timed("grow", self._grow, 1, e, env_bs, t_e, label, label)
            self.sys_trun[label + 1], self.env_trun[label + 1] = t_s, t_e
            self.sys_basis[label + 1], self.env_basis[label + 1] = sys_bs, env_bs
            label += 1

    # ---- left_to_right (dmrg_sub.f90:728-907) -----------------------------
    def left_to_right(self, bg_id, end_id):
        N = self.nsite
        label = bg_id
        trun = None
        while label + 1 <= end_id:
            s, e = self.sys[label], self.env[N - label - 2]
            sys_bs = self._basis(s)
            env_bs = self.env_basis.get(N - label - 1) or self._basis(e)
            wsec = get_super_basis(sys_bs, env_bs, self.total_up, 0)
            psi0 = None
            if self.wsec is not None and np.array_equal(self.wsec, wsec):
                psi0 = self.psi               # the same bond as the last solve
            elif trun is not None:
                # wave_trans_LR(sys_bs, env_basis, sys_trun, env_trun, old, new)
                psi0 = self._timed("transform", self._transform, 0, sys_bs, self.env_basis[N - label], trun,
                                   self.env_trun[N - label - 1], wsec)
            info, secs = self._solve(s, e, sys_bs, env_bs, wsec, psi0, "left-to-right")
            trun, kept, err = self._timed("truncate", self._truncate, 0, label, N - label - 2, wsec, sys_bs, self.psi,
                                          False)
            self._record("left-to-right", label, N - label - 2, wsec, info, kept, err, secs)
            self.sys[label + 1] = self._timed("grow", self._grow, 0, s, sys_bs, trun, label, N - label - 2)
            self.sys_trun[label + 1], self.sys_basis[label + 1] = trun, sys_bs
            label += 1

    # ---- right_to_left (dmrg_sub.f90:910-1087) ----------------------------
    def right_to_left(self, bg_id, end_id):
        N = self.nsite
        label = bg_id
        trun = None
        while label - 1 >= end_id:
            s, e = self.sys[label - 1], self.env[N - label - 1]
            sys_bs = self.sys_basis.get(label) or self._basis(s)
            env_bs = self._basis(e)
            wsec = get_super_basis(sys_bs, env_bs, self.total_up, 0)
            psi0 = None
            if self.wsec is not None and np.array_equal(self.wsec, wsec):
                psi0 = self.psi
            elif trun is not None:
                # wave_trans_RL(sys_basis, env_bs, sys_trun, env_trun, old, new)
                psi0 = self._timed("transform", self._transform, 1, self.sys_basis[label + 1], env_bs,
                                   self.sys_trun[label], trun, wsec)
            info, secs = self._solve(s, e, sys_bs, env_bs, wsec, psi0, "right-to-left")
            trun, kept, err = self._timed("truncate", self._truncate, 1, label - 1, N - label - 1, wsec, env_bs,
                                          self.psi, False)
            self._record("right-to-left", label - 1, N - label - 1, wsec, info, kept, err, secs)
            self.env[N - label] = self._timed("grow", self._grow, 1, e, env_bs, trun, label - 1, N - label - 1)
            self.env_trun[N - label], self.env_basis[N - label] = trun, env_bs
            label -= 1

    # ---- checkpoint files (block.f90:2760-3097) ---------------------------
    @staticmethod
    def _host(dev) -> OpSet:
        n = int(dev.ops.data.size)
        return OpSet(dev.ops.length, dev.ops.hdr, dev.ops.sub, dev.ops.off, dev.data[:n].cpu().numpy())

    def save_checkpoint(self, directory):
        """Write what the run holds as the reference's checkpoint files:
        sysblocks<L> / envblocks<L> for every block model, systruns<L> /
        envtruns<L> for the truncation that made block L, sysbs<L> / envbs<L>
        for the fused basis (block L-1 + site), and eigvec1 for the ground
        state of the last bond (model_to_disk, truns_to_disk, basis_to_disk,
        wave_to_disk, block.f90:2760-3045; names of Get_model_name,
        dmrg_sub.f90:294).  The reference, or load_checkpoint, resumes from
        them."""
        from . import checkpoint as ck
        os.makedirs(directory, exist_ok=True)
        sides = ((0, self.sys, self.sys_trun, self.sys_basis), (1, self.env, self.env_trun, self.env_basis))
        for side, blocks, truns, bases in sides:
            for L, blk in blocks.items():
                ck.save_model(os.path.join(directory, ck.file_name(side, "blocks", L)), self._host(blk))
            for L, t in truns.items():
                ck.save_trun(os.path.join(directory, ck.file_name(side, "truns", L)), self._host(t))
            for L, b in bases.items():
                ck.save_basis(os.path.join(directory, ck.file_name(side, "bs", L)), b)
        if self.psi is not None:
            ck.save_wave(os.path.join(directory, ck.wave_file(1)), self.bond[0] + 1, self.bond[1] + 1, self.wsec,
                         self.psi.cpu().numpy())

    def load_checkpoint(self, directory):
        """Read checkpoint files written by save_checkpoint or by the
        reference (model_from_disk, truns_from_disk, basis_from_disk,
        wave_from_disk) into the run: block models and truncation operators
        go to device memory, bases stay on the host, and eigvec1 becomes the
        start vector of the next solve at its bond.  A sweep
        (left_to_right / right_to_left) then continues from them without the
        warm-up."""
        from . import checkpoint as ck
        pat = re.compile(r"^(sys|env)(blocks|truns|bs)([0-9]+)$")
        for name in sorted(os.listdir(directory)):
            m = pat.match(name)
            if not m:
                continue
            side, kind, L = 0 if m.group(1) == "sys" else 1, m.group(2), int(m.group(3))
            path = os.path.join(directory, name)
            if kind == "blocks":
                ops, _ = ck.load_model(path)
                if ops.length != L:
                    raise ck.CheckpointError(f"{path}: holds a block of length {ops.length}")
                (self.sys if side == 0 else self.env)[L] = self._block(ops)
            elif kind == "truns":
                t, _ = ck.load_trun(path)
                trun = DeviceTrun(OpSet(L, t.hdr, t.sub, t.off, np.zeros(max(1, t.data.size))),
                                  self._device(t.data if t.data.size else np.zeros(1)))
                (self.sys_trun if side == 0 else self.env_trun)[L] = trun
            else:
                b, _ = ck.load_basis(path)
                (self.sys_basis if side == 0 else self.env_basis)[L] = b
        wave = os.path.join(directory, ck.wave_file(1))
        if os.path.exists(wave):
            w = ck.load_wave(wave)
            self.psi, self.wsec = self._device(w.x), w.wsec
            self.bond = (w.sys_len - 1, w.env_len - 1)

    # ---- measurements (measure_operator_dia, measure.f90:3-57) -----------
    def measure_density(self):
        """<n_i> of every lattice site in the ground state of the last bond,
        carried through the run's bases and truncation operators."""
        from . import measure
        if self.psi is None:
            raise RuntimeError("no ground state yet")
        sl, el = self.bond[0] + 1, self.bond[1] + 1
        return measure.density_profile(self.model.up_bias, sl, el, self.sys_basis, self.env_basis, self.sys_trun,
                                       self.env_trun, self.wsec, self.psi)

    def measure_cross_correlation(self):
        """<n_i n_j> for the system-side sites i (lattice 1..sys_len) and the
        environment-side sites j (lattice N - j + 1) of the last bond."""
        from . import measure
        if self.psi is None:
            raise RuntimeError("no ground state yet")
        sl, el = self.bond[0] + 1, self.bond[1] + 1
        return measure.cross_correlation(self.model.up_bias, sl, el, self.sys_basis, self.env_basis, self.sys_trun,
                                         self.env_trun, self.wsec, self.psi)

    def measure_correlations(self):
        """The whole <n_i n_j> matrix (lattice order) in the ground state of
        the last bond (measure.correlation_matrix)."""
        from . import measure
        if self.psi is None:
            raise RuntimeError("no ground state yet")
        sl, el = self.bond[0] + 1, self.bond[1] + 1
        return measure.correlation_matrix(self.model.up_bias, sl, el, self.sys_basis, self.env_basis, self.sys_trun,
                                          self.env_trun, self.wsec, self.psi)

    def measure_pair_correlations(self, kind, sid1):
        """Get_density_cor's two-point functions (measure.pair_correlations)
        of lattice site sid1 on the system side with every later site, in
        the ground state of the last bond: kind "density" <n_i n_j>, "spin"
        <S_i . S_j>, "hopping" sum over s of <c^+_is c_js>."""
        from . import measure
        if self.psi is None:
            raise RuntimeError("no ground state yet")
        sl, el = self.bond[0] + 1, self.bond[1] + 1
        return measure.pair_correlations(kind, self.model.up_bias, sid1, sl, el, self.sys_basis, self.env_basis,
                                         self.sys_trun, self.env_trun, self.wsec, self.psi)

    def measure_line_correlations(self, kind, xi, xj, ij=0):
        """Get_density_cor's matrix oper_ndia (measure.density_cor) for the
        sites xi..xj: the pairs along the lattice's legs, as Measurement
        writes them to density_cor.dat, spin_cor.dat and elec_cor.dat
        (main.f90:272-318)."""
        from . import measure
        if self.psi is None:
            raise RuntimeError("no ground state yet")
        sl, el = self.bond[0] + 1, self.bond[1] + 1
        return measure.density_cor(kind, self.model.up_bias, xi, xj, self.model.ny, ij, sl, el, self.sys_basis,
                                   self.env_basis, self.sys_trun, self.env_trun, self.wsec, self.psi)

    def _neighbours(self):
        from .lattice import square_neighbours
        m = self.model
        return square_neighbours(m.ny, m.nx, m.nsite)

    def measure_sc_order(self, xc1, xc2):
        """Get_SC_Order (measure.sc_order) in the ground state of the last
        bond: the singlet pair amplitude of the y and x bonds of the sites
        xc1, xc1 + ny, ... <= xc2, as Measurement writes Sc_order.dat
        (main.f90:243-246 with xc1 = 1 and 2, xc2 = N - ny).  Needs the
        SU(2)-only basis (up_bias 0)."""
        from . import measure
        if self.psi is None:
            raise RuntimeError("no ground state yet")
        sl, el = self.bond[0] + 1, self.bond[1] + 1
        return measure.sc_order(self.model.up_bias, self._neighbours(), self.model.ny, xc1, xc2, sl, el,
                                self.sys_basis, self.env_basis, self.sys_trun, self.env_trun, self.wsec, self.psi)

    def measure_sc_pair_cor0(self, xc1, ij):
        """Get_mSCOP_Operator_Cor0 (measure.sc_pair_cor0) in the ground state
        of the last bond: the pair-pair correlations of the bond of site xc1
        in direction ij (1: y, 3: x) with the bonds further along, as
        Measurement writes Sc_cor_y.dat / Sc_cor_x.dat (main.f90:255-259)."""
        from . import measure
        if self.psi is None:
            raise RuntimeError("no ground state yet")
        sl, el = self.bond[0] + 1, self.bond[1] + 1
        return measure.sc_pair_cor0(self.model.up_bias, self._neighbours(), self.model.ny, xc1, ij, sl, el,
                                    self.sys_basis, self.env_basis, self.sys_trun, self.env_trun, self.wsec, self.psi)

    def ground_state(self, lanerr, trun_err, kept_max, kept_maxf, sweep_num, nsite0=None):
        """Get_Ground_State (main.f90:96-187) for a fresh run (restart1 =
        restart2 = 0, iter1 = 0), with the kept-state settings of main.f90:
        13-33 (kept_min = min(1000, 3 kept_max / 4), kept_max capped at 1200).
        nsite0: Num_site0 of the reference, Nleg (Nx + 2) (default), which
        sets where the closing left-to-right pass ends.  Returns the last
        ground-state energy."""
        N = self.nsite
        if nsite0 is None:
            nsite0 = self.model.ny * (self.model.nx + 2)
        self.kept_max = min(1200, kept_max)
        self.kept_min = min(1000, kept_max * 3 // 4)
        # warm-up
        self.tol, self.trun_err = lanerr * 1000, trun_err * 200
        self.warm_up()
        self.tol, self.trun_err = lanerr * 30, trun_err * 15
        self.kept_max = self.kept_max * 2 // 3 + kept_maxf // 3
        self.kept_min = min(3 * self.kept_max // 4, 3600)
        # the first half sweep starts from a random vector, as left_to_right
        # does with restart1 = 0 (dmrg_sub.f90:763-766)
        self.wsec = None
        self.left_to_right(N // 2 - 1, N - 2)
        self.kept_min = min(3 * self.kept_max // 4, 3600)
        self.right_to_left(N - 2, 1)
        self.kept_max = kept_maxf
        self.kept_min = min(3 * self.kept_max // 4, 3600)
        self.tol, self.trun_err = lanerr, trun_err
        for _ in range(sweep_num):
            self.left_to_right(1, N - 2)
            self.right_to_left(N - 2, 1)
        self.left_to_right(1, nsite0 // 2 + 3)
        return self.log[-1].energy

    def run(self, sweeps=2):
        """Warm-up, then `sweeps` pairs of half sweeps (Get_Ground_State with
        trun_idx = 1).  Returns the last ground-state energy."""
        N = self.nsite
        self.warm_up()
        self.left_to_right(N // 2 - 1, N - 2)
        self.right_to_left(N - 2, 1)
        for _ in range(sweeps):
            self.left_to_right(1, N - 2)
            self.right_to_left(N - 2, 1)
        return self.log[-1].energy
