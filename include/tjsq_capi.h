/*
 * tjsq_capi.h -- C-ABI of the MI355X tjsq Lanczos / H_eff library
 * (libtjsq_hip.so, built from grand_sql_tj_amd/csrc).
 *
 * The library replaces the Lanczos step of the reference tjsq SU(2) two-site
 * DMRG code (DongNingSheng/Grand_SQL_tJ):
 *
 *   lanczos_memo_new(sys,env,sys_bs,env_bs,super,eigvec,levels,num_idx)
 *                                          code/tjsq/eigen.f90:2-10
 *   hmerge(sys,env,sys_bs,env_bs)          code/tjsq/get_sys.f90:4383-4387
 *   model_wave(sys,env,sys_bs,env_bs,invec,outvec)
 *                                          code/tjsq/get_sys.f90:4127-4134
 *
 * and the fusion-table routines they rely on
 *
 *   Get_basis(block,site,basis)            code/tjsq/block.f90:9-104
 *   Get_super_basis(sys_bs,env_bs,super,total_up,total_down)
 *                                          code/tjsq/block.f90:110-161
 *   W6JS(j1,j2,j3,l1,l2,l3)                code/tjsq/block.f90:1564-1615
 *
 * The reference's derived types are not C-interoperable; every table crosses
 * this boundary flattened (layout below).  There are no hidden globals: the
 * lattice couplings that the reference keeps in module pubdata
 * (code/tjsq/common.f90:35-45) arrive in tjsq_lattice.  Every call returns an
 * int status (TJSQ_OK = 0); the reference stops the program instead
 * (get_sys.f90:2042-2045 and similar), the Fortran shim turns a nonzero
 * status into `error stop`.
 *
 * Flat layout (all integer tables int32, column-major matrices):
 *   operator set (Total_Model, common.f90:139-149): nops operators; op 0 is
 *     the block Hamiltonian, op 1+4*(i-1)+q is active site i's elec_up (q=0),
 *     elec_down (1), spin_sd (2), num_sn (3).
 *     hdr[4*op]  = {len, up_dif, down_dif, nsub}
 *     sub[5*s]   = {num_up, num_down, down_dif, row_dim, sdim}  (ket QNs)
 *     off[s]     = element offset of sub-block s inside data
 *   basis (Total_Basis, common.f90:65-82):
 *     sec[4*k]   = {new_num_up, new_num_down, dim, nsub}
 *     sub[6*q]   = {bl_num_up, bl_num_down, st_num_up, st_num_down, spos, sdim}
 *   wavefunction sectors (Super_Basis, common.f90:88-97):
 *     wsec[6*i]  = {sys_num_up, sys_num_down, env_num_up, env_num_down,
 *                   sys_dim, env_dim}; a vector is the sectors back to back,
 *     each column-major sys_dim x env_dim.
 *
 * Threading: a plan is used by one host thread and one stream at a time;
 * distinct plans are independent.
 */
#ifndef TJSQ_CAPI_H
#define TJSQ_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TJSQ_OK 0
#define TJSQ_ERR_ARG (-1)     /* invalid argument or inconsistent tables */
#define TJSQ_ERR_QN (-2)      /* quantum-number mismatch between operators */
#define TJSQ_ERR_HIP (-3)     /* HIP runtime error */
#define TJSQ_ERR_NOMEM (-4)   /* device or host allocation failed */
#define TJSQ_ERR_NODEV (-5)   /* no usable GPU */
#define TJSQ_ERR_CAPACITY (-6)/* request exceeds admitted memory */

#define TJSQ_FP64 0           /* FP64 MFMA contractions */
#define TJSQ_FP32 1           /* FP32 MFMA contractions, FP32 accumulation */
#define TJSQ_BF16X3 2         /* BF16 hi/lo split, 3 MFMA products, FP32 acc. */

/* Lattice and couplings in the reference's conventions (main.f90:35-47):
 * jt = -sqrt(2) t, jd = -sqrt(3) J, jn = -J/4 for nearest-neighbour bonds
 * (index 0) and next-nearest bonds (index 1); esite = -mu on every site.
 * Square cylinder, periodic in y (ny = circumference), open in x; the
 * active-site tables (dmrg_setup0) are built over nx_tables columns. */
typedef struct tjsq_lattice {
  int ny;
  int nx_tables;
  double jt[2];
  double jd[2];
  double jn[2];
  double esite;
  int up_bias;   /* 1: sectors carry (N, 2S); 0: 2S only (reference default) */
} tjsq_lattice;

typedef struct tjsq_opset {
  int length;            /* block length (Total_Model%len) */
  int nops;
  const int* hdr;
  const int* sub;
  const int64_t* off;
  const double* data;    /* host or device pointer, see data_on_device */
  int64_t ndata;
  int data_on_device;    /* 1: data is a device pointer (used in place) */
} tjsq_opset;

typedef struct tjsq_basis {
  int length;
  int nsec;
  const int* sec;
  const int* sub;
} tjsq_basis;

typedef struct tjsq_plan tjsq_plan;

typedef struct tjsq_plan_stats {
  int64_t dim;              /* D, length of the wavefunction */
  int64_t n_contrib;        /* GEMM-chain / GEMM / axpy contributions */
  int64_t n_chain;          /* two-sided contributions (A.X.B) */
  int64_t n_stage1;         /* distinct A.X products of stage 1 */
  int64_t n_out_blocks;     /* output sub-blocks */
  int64_t n_tiles1, n_tiles2;
  double flops_ref;         /* algorithmic flops per apply, reference order */
  double flops_exec;        /* flops the two stages execute */
  double bytes_fp64;        /* compulsory bytes per apply at 8 B/element */
  int64_t op_elems;         /* operator elements referenced */
  int64_t scratch_elems;    /* stage-1 scratch elements */
  int64_t n_slots;          /* partial output tiles of split stage-2 work */
  int64_t packed;           /* 1: the BF16x3 apply runs on the packed layout
                               (tjsq_hostplan_packed), 0: on operand words */
  double flops_packed;      /* MFMA flops per product the packed tiles execute
                               (padding included; BF16x3 issues 3 products) */
} tjsq_plan_stats;

typedef struct tjsq_lanczos_opts {
  double tol;          /* |E0(k) - E0(k-1)| stop (lanzero, eigen.f90:98) */
  int max_iter;        /* loop_in (eigen.f90:12) */
  int precision;       /* TJSQ_FP64 / TJSQ_FP32 / TJSQ_BF16X3 */
  double refine_tol;   /* FP64 residual |H psi - E psi| above which a mixed
                          solve restarts in FP64 from its own result */
  int max_refine;      /* FP64 restarts allowed after a mixed solve */
  double energy_tol;   /* > 0: a mixed solve restarts in FP64 only when the
                          estimated energy error |r|^2 / (theta_2 - E) of its
                          FP64 Rayleigh quotient exceeds energy_tol (theta_2:
                          second Ritz value); refine_tol is then unused */
  double rtol;         /* relative Ritz-value change that also ends a pass:
                          |dE| <= max(tol, rtol |E|); 0 selects 1e-7 for
                          mixed passes (their rounding floor) and none for
                          FP64 passes */
} tjsq_lanczos_opts;

typedef struct tjsq_lanczos_result {
  double e0;           /* ground-state energy (FP64 Rayleigh quotient for
                          mixed precision, Ritz value for FP64) */
  double e_ritz;       /* last Ritz value of the tridiagonal problem */
  double residual;     /* FP64 |H psi - e0 psi| (0 if not computed) */
  int iterations;      /* Lanczos vectors built, all passes */
  int applies;         /* H_eff applies, all precisions */
  int refinements;     /* FP64 restarts taken */
  int breakdown;       /* 1 if beta <= tol ended a pass (eigen.f90:74-77) */
  int refine_iterations; /* Lanczos vectors of the FP64 restarts */
  double gap;          /* theta_2 - E of the last pass (0 if unknown) */
  double energy_err_est; /* |r|^2 / gap of the returned state */
  double seconds;      /* wall time of the solve */
} tjsq_lanczos_result;

/* ---- host tables, bit-exact with the reference ------------------------ */

/* 6j symbol on doubled arguments (W6JS). */
double tjsq_w6js(int j1, int j2, int j3, int l1, int l2, int l3);
/* 3j symbol (j1 j2 j3; m1 m2 m3), all arguments twice the angular momentum
 * (W3JS, code/tjsq/block.f90:1522-1560). */
double tjsq_w3js(int j1, int j2, int j3, int m1, int m2, int m3);

/* Get_basis: fuse a block with a site.  blk[3*i] = {num_up, num_down, dim}
 * of the block's sectors in order, site[3*j] likewise.  sec must hold
 * 3*nb*nsite*4 ints and sub 3*nb*nsite*6 ints. */
int tjsq_get_basis(int block_len, int nb, const int* blk, int nsite, const int* site,
                   int* sec, int* sub, int* nsec, int* nsub);

/* Get_super_basis: wavefunction sectors for total (N, 2S).  wsec must hold
 * 6*sys->nsec*env->nsec ints. */
int tjsq_get_super_basis(const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int total_up,
                         int total_down, int* wsec, int* nw);

/* Coupling structure of hmerge (coup_bs[4] = (1,1),(1,2),(2,1),(2,2),
 * coup_bb[2], inb[0..sys_len+env_len+2]). */
int tjsq_hmerge_tables(const tjsq_lattice* lat, int sys_len, int env_len, int* coup_bs,
                       int* coup_bb, int* inb);

/* Active-site count nleg11(len) of dmrg_setup0. */
int tjsq_active_sites(const tjsq_lattice* lat, int len);

/* ---- plans and applies ------------------------------------------------ */

/* Compile hmerge + model_wave for one superblock into a device plan:
 * uploads (or adopts) the operator data, builds the merged operators on the
 * device and the grouped-GEMM task tables.  precision_mask selects which
 * operand copies to prepare (bit p for TJSQ_FP64/FP32/BF16X3); FP64 is always
 * prepared (its stage-1 scratch and partial tiles are allocated on the first
 * FP64 apply when bit 0 is clear).  stream may be NULL (default stream). */
int tjsq_plan_create(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys,
                     const tjsq_opset* env, const tjsq_basis* sys_bs, const tjsq_basis* env_bs,
                     int nw, const int* wsec, int precision_mask, void* stream, tjsq_plan** plan);

/* Work-item shape of the grouped GEMM (what an autotuner varies).  Blocks
 * with a side above `small` use big x big work items (big = 64 or 32;
 * small = 32); stage-2 segments are cut into K pieces of at most k_split and
 * grouped into work items of at most item_chunks K chunks of 32; work items
 * are binned over `xcds` XCDs (1 = plain order).  merge_right = 1 factors
 * the two-sided terms of one output block that share their right operand B:
 * sum_j c_j A_j X_j B = (sum_j c_j A_j X_j) B, the sum formed in stage 1.
 * order = 1 lays stage-2 work items out by column strip (output blocks that
 * share an env sub-block, hence their env operators) with each block's
 * segments sorted by right operand; 0 keeps wave-sector order.
 * NULL selects {64, 32, 256, 32, 1, 1, 0}. */
typedef struct tjsq_shape {
  int big;
  int small;
  int k_split;
  int item_chunks;
  int xcds;
  int merge_right;
  int order;
} tjsq_shape;

int tjsq_plan_create_ex(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys,
                        const tjsq_opset* env, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int nw,
                        const int* wsec, int precision_mask, const tjsq_shape* shape, void* stream,
                        tjsq_plan** plan);

/* Recompile a plan's task tables for another work-item shape, keeping its
 * operators on the device (no launch of the plan may be in flight). */
int tjsq_plan_reshape(tjsq_plan* plan, const tjsq_shape* shape, void* stream);

/* New operator data for a plan whose block layouts are unchanged (the same
 * hdr / sub / off of both sets as at tjsq_plan_create; TJSQ_ERR_ARG
 * otherwise): the data are uploaded, hmerge is run again and the FP32 /
 * BF16x3 copies are rebuilt, with no host compile.  No reference
 * counterpart: the reference's model_wave and lanczos_memo_new
 * (code/tjsq/get_sys.f90:4127-4134, code/tjsq/eigen.f90:2-10) read the block
 * models on every call; this lets a caller that holds a plan per bond
 * follow the models as a sweep rebuilds them (the Fortran drop-in's plan
 * cache). */
int tjsq_plan_set_ops(tjsq_plan* plan, const tjsq_opset* sys, const tjsq_opset* env, void* stream);

int tjsq_plan_stats_get(const tjsq_plan* plan, tjsq_plan_stats* stats);
int tjsq_plan_destroy(tjsq_plan* plan);

/* y = H_eff x on device FP64 vectors of length D (model_wave). */
int tjsq_heff_apply(tjsq_plan* plan, const double* x, double* y, void* stream);

/* y = H_eff x on device FP32 vectors, contractions in TJSQ_FP32 or
 * TJSQ_BF16X3. */
int tjsq_heff_apply_f32(tjsq_plan* plan, int precision, const float* x, float* y, void* stream);

/* tjsq_heff_apply (precision TJSQ_FP64, double vectors) or
 * tjsq_heff_apply_f32 with TJSQ_FP32 (float vectors), allocating nothing:
 * the stage-1 products and partial tiles live in a static device array of
 * TJSQ_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  Plans
 * needing more scratch than the array holds are rejected (TJSQ_ERR_ARG), as
 * are BF16x3 and a plan with profiling on.  Calls that overlap on one GPU
 * share the array and race.  Results are bitwise those of the allocating
 * entry points. */
#define TJSQ_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int tjsq_heff_apply_noalloc(tjsq_plan* plan, int precision, const void* x, void* y, void* stream);

/* Per-launch timing of the apply kernels with HIP events recorded on the
 * launching stream.  profile_read returns the summed milliseconds and launch
 * counts since the last read in ms[5] / launches[5]: stage-1 64x64 items,
 * stage-1 32x32 items, stage-2 64x64 items, stage-2 32x32 items, and the
 * partial-tile reduction.  The stages are distinct kernel symbols
 * (k_tiles<mode, edge, stage>), so a rocprofv3 trace reports them apart. */
int tjsq_plan_profile(tjsq_plan* plan, int enable);
int tjsq_plan_profile_read(tjsq_plan* plan, double* ms, int64_t* launches);

/* Capture each Lanczos iteration (recurrence, normalisation, Gram-Schmidt,
 * apply, alpha) as a hipGraph on its first use and replay it in later
 * iterations and solves of this plan (enable = 1).  Needs a non-default
 * stream; results are bitwise those of the launched form.  One host read of
 * (beta, alpha) per iteration remains for the tridiagonal problem. */
int tjsq_plan_graphs(tjsq_plan* plan, int enable);

/* Ground state (lanczos_memo_new, levels = 1): psi (device, FP64, length D)
 * holds the start vector on entry and the normalised ground state on exit. */
int tjsq_lanczos_solve(tjsq_plan* plan, double* psi, const tjsq_lanczos_opts* opts,
                       tjsq_lanczos_result* result, void* stream);

/* ---- host-only plan inspection (no GPU needed) --------------------------
 * The task tables a plan would upload, for tests of the plan compiler.
 * sizes[15] = {n_seg1, n_tile1, n_seg2, n_tile2, n_merge, arena_elems,
 * scratch_elems, dim, sys_base, env_base, merged_base, n_rtile, n_slots,
 * n64_1, n64_2}; the first n64_s work items of stage s are 64x64 tiles, the
 * rest 32x32.
 * Table element layouts: segment 40 bytes {int64 l_off, r_off; int32 l_ld,
 * r_ld, K; uint8 l_space, r_space, l_t, r_t; double c}; tile 48 bytes
 * {int64 out_off; int32 out_ld, m0, n0, mt, nt, seg_begin, seg_end,
 * out_space, slot, group}; merge term 32 bytes {int64 dst_off, src_off, n; double
 * coef}.  A stage-2 tile with slot >= 0 writes its partial tile (leading
 * dimension = its edge) at element 1024 * slot; reduce tiles (rtile, tile
 * layout) sum the partials at units [seg_begin, seg_end) in steps of `slot`
 * units, leading dimension out_space, in order into their output tile.
 * Work items are ordered for the round-robin XCD dispatch (item g runs on
 * XCD g % 8) so that items sharing operands (same `group`) share one L2;
 * padding items have mt = nt = 0 and no segments. */
typedef struct tjsq_hostplan tjsq_hostplan;
int tjsq_hostplan_create(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys,
                         const tjsq_opset* env, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int nw,
                         const int* wsec, tjsq_hostplan** out);
int tjsq_hostplan_create_ex(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys,
                            const tjsq_opset* env, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int nw,
                            const int* wsec, const tjsq_shape* shape, tjsq_hostplan** out);
int tjsq_hostplan_sizes(const tjsq_hostplan* h, int64_t* sizes);
int tjsq_hostplan_tables(const tjsq_hostplan* h, void* seg1, void* tile1, void* seg2, void* tile2, void* merge,
                         void* rtile);
int tjsq_hostplan_stats(const tjsq_hostplan* h, tjsq_plan_stats* stats);
int tjsq_hostplan_destroy(tjsq_hostplan* h);
/* The packed tables of the BF16x3 apply (the layout the plan uses for
 * BF16x3 when it applies): operands as lines of k in groups of eight (eight
 * bf16 high halves, then eight low halves), every sub-block zero-padded to
 * multiples of 16.  sizes[10] = {n_ops_jobs, n_x_jobs, n_seg1, n_tile1,
 * n_seg2, n_tile2, n_rtile, n_slots, ops_bytes, x_bytes} plus sizes[10] =
 * t_bytes; then the tables are copied to the non-null pointers.  Layouts:
 * pack job 40 bytes {int64 src_off, dst_off; int32 src_ld, rows, cols,
 * trans, lines, kpad}; segment 48 bytes {int64 l_off, r_off; uint32 l_lim,
 * r_lim; int32 l_stride, r_stride, K, spaces; double c}; tile 56 bytes
 * {int64 out_off; int32 out_ld, m0, n0, mb, nb, mlim, nlim, seg_begin,
 * seg_end, out_kind, slot, pad}; reduce 40 bytes {int64 out_off; int32
 * out_ld, m0, n0, mt, nt, s0, s1, pad}.  Returns TJSQ_ERR_ARG where the
 * layout does not apply (the plan then keeps the unpacked BF16x3 apply).
 * No counterpart in the reference: test-side inspection of the tables. */
int tjsq_hostplan_packed(const tjsq_hostplan* h, int64_t* sizes, void* ops_jobs, void* x_jobs, void* seg1,
                         void* tile1, void* seg2, void* tile2, void* rtile);

/* ---- host-buffer entry points (Fortran callers holding data in host
 * memory): the same operations with x, y, psi in host memory; the library
 * stages them through device buffers of the plan (PCIe transfers included). */
int tjsq_heff_apply_host(tjsq_plan* plan, const double* x, double* y);
int tjsq_lanczos_solve_host(tjsq_plan* plan, double* psi, const tjsq_lanczos_opts* opts,
                            tjsq_lanczos_result* result);

/* ---- block update: density matrix, truncation, enlargement ------------
 * After a solve the reference forms the reduced density matrix, keeps its
 * largest eigenvectors and enlarges/projects the block operators
 * (code/tjsq/dmrg_sub.f90:656-693).  These calls do the same on the device.
 *
 * A program is compiled on the host by a *_create call (no GPU needed) and
 * run with tjsq_program_run(prog, ops, x, y, stream) on device pointers:
 *   block update / projection: ops = truncation-operator data, x = the old
 *     block's operator data, y = the new operator data (layout from
 *     tjsq_program_layout);
 *   density matrix: x = psi (levels states of length D back to back), ops
 *     unused, y = rho.
 * Every output sub-block is written in full.  sizes[14] = {nops, nsub,
 * ndata, x_elems, ops_elems, n_seg1, n_tile1, n_seg2, n_tile2, n_rtile,
 * n_slots, scratch_elems, n64_1, n64_2}; the task tables have the layouts
 * of tjsq_hostplan_tables, with space codes 0 = ops, 1 = x, 2 = stage-1
 * scratch, 3 = identity. */
typedef struct tjsq_program tjsq_program;

/* Get_Hamiltonian_Sys (side 0) / Get_Hamiltonian_Env (side 1)
 * (code/tjsq/get_sys.f90:4-246, 247-478): enlarge the block model `pre`
 * (length L = sys_len or env_len, nops = 1 + 4 * active sites) by one site
 * over the fused basis `basis` (Get_basis of the block with the site) and
 * project every operator with the truncation operator `trun` (one operator
 * over the basis sectors).  The result is the model at length L + 1 with
 * nleg11(L + 1) active sites, in the operator-set layout. */
int tjsq_block_update_create(const tjsq_lattice* lat, int side, int sys_len, int env_len, const tjsq_opset* pre,
                             const tjsq_basis* basis, const tjsq_opset* trun, tjsq_program** prog);

/* update_trun_dia (ndia[i] = 0) / update_trun_ndia (ndia[i] = 1) of every
 * operator of ops with one truncation operator (code/tjsq/block.f90:
 * 1311-1503); ndia may be NULL (all diagonal). */
int tjsq_trun_project_create(const tjsq_opset* ops, const int* ndia, const tjsq_opset* trun, tjsq_program** prog);

/* One step of the transport of diagonal operators along a block
 * (Get_sys_oper_dia / Get_env_oper_dia, code/tjsq/get_sys.f90:764-891, the
 * operator transport of the measurements, code/tjsq/measure.f90:3-57): every
 * operator of `ops` (diagonal, at block length L; may be NULL) embedded into
 * the fused basis `basis` of length L + 1 (update_block_dia), then the new
 * site's operator `site` (one operator of 1x1 sub-blocks, host data; may be
 * NULL) embedded (update_site_dia), each projected with `trun`
 * (update_trun_dia).  Run with ops = trun data, x = the data of `ops`. */
int tjsq_dia_transport_create(const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis* basis,
                              const tjsq_opset* trun, tjsq_program** prog);

/* tjsq_dia_transport_create with the products of the first nprod
 * operators of `ops` with the site operator appended to the output (output
 * order for nprod > 0: ops[0, nprod), the site operator, ops[nprod, end),
 * the products), each product a reduced operator of the fused basis
 * (sqrt((1+2S)/(1+2S_bl)) A(bl) times the site operator's value
 * s(st)/sqrt(1+2S_st)), then projected: the two-point operators n_i n_j
 * of a block, formed when site j joins it.  No reference counterpart: the
 * reference forms such products with block_mul_block_dia
 * (code/tjsq/block.f90:1924-1954), which multiplies reduced matrices
 * without the SU(2) factor. */
int tjsq_dia_transport_pairs_create(const tjsq_opset* ops, int nprod, const tjsq_opset* site, const tjsq_basis* basis,
                                    const tjsq_opset* trun, tjsq_program** prog);

/* Expectation values in a wavefunction of diagonal block operators and of
 * the site operator next to the block (measure_sys_block / measure_sys_site
 * for side 0, measure_env_block / measure_env_site for side 1,
 * code/tjsq/measure.f90:1241-1470).  `basis` is the fused basis of the
 * measured side, wsec the wavefunction's sector list.  The output is
 * nops + 1 vectors of the wavefunction's length, one per operator of `ops`
 * and the site operator's last: <O_j> = psi . y_j.  Run with ops = the
 * operators' data, x = psi, and y cleared first (entries no term reaches
 * are not written). */
int tjsq_measure_create(int side, const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis* basis, int nw,
                        const int* wsec, tjsq_program** prog);

/* One step of the transport of rank-k operators along a block (the steps of
 * Get_sys_oper_ndia / Get_env_oper_ndia and Change_sys_oper_ndia /
 * Change_env_oper_ndia, code/tjsq/get_sys.f90:998-1238, that Get_density_cor,
 * code/tjsq/measure.f90:3110-3223, runs per site): every operator of `ops`
 * (at block length L; may be NULL) embedded into the fused basis `basis` of
 * length L + 1 (update_block_ndia_sys for side 0, update_block_ndia_env for
 * side 1, code/tjsq/block.f90:503-920), then the new site's operator `site`
 * (one operator of 1x1 sub-blocks, host data; may be NULL) embedded
 * (update_site_ndia_sys / _env, :925-1305), each projected with `trun`
 * (update_trun_ndia, :1360-1503).  flag is 'B' or 'F', the reference's
 * Types argument (fermion signs of the embeddings).  Run with ops = trun
 * data, x = the data of `ops`. */
int tjsq_ndia_transport_create(int side, int flag, const tjsq_opset* ops, const tjsq_opset* site,
                               const tjsq_basis* basis, const tjsq_opset* trun, tjsq_program** prog);

/* The same-side two-site operators of Get_sys2b (side 0) / Get_env2b
 * (side 1) (code/tjsq/measure.f90:2748-2819) at the block length L where the
 * second site joins: each operator of `ops` (rank k, block length L - 1)
 * embedded into `basis` (update_block_ndia_*), the site operator `site`
 * embedded (update_site_ndia_*), their rank-0 coupling (block_mul_block_ndia
 * with down_dif 0, code/tjsq/block.f90:1956-2197), projected with `trun`
 * (update_trun_dia).  Output: one diagonal operator at length L per operator
 * of `ops`, to be carried on with tjsq_dia_transport_create and measured
 * with tjsq_measure_create.  Run with ops = trun data, x = the data of
 * `ops`. */
int tjsq_pair_ndia_create(int side, int flag, const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis* basis,
                          const tjsq_opset* trun, tjsq_program** prog);

/* The across-the-bond correlations of Get_density_cor (code/tjsq/measure.f90:
 * 3110-3223) for one system block operator `sys_op` (block length
 * sys_len - 1): with every operator of `env_ops` (environment block length
 * env_len - 1; may be NULL; sys_block_env_block_cor_ndia, :5241-5480), with
 * the environment bond-site operator `env_site` (may be NULL;
 * sys_block_env_site_cor_ndia, :5482-5706) and with the system bond-site
 * operator `sys_site` (may be NULL; sys_block_site_cor_ndia, :4939-5083).
 * *_trans != 0 applies block_transfer_trans (code/tjsq/block.f90:1676-1706)
 * to that operand first, as Get_density_cor does.  sys_basis / env_basis are
 * the wavefunction's fused bases (sys_bsm(sys_len), env_bsm(env_len)), wsec
 * its sector list.  The output is one vector of the wavefunction's length per
 * operator of env_ops, then env_site's, then sys_site's: each routine's
 * value = psi . y_q.  Run with ops = the data of sys_op followed by the data
 * of env_ops, x = psi, y cleared first. */
int tjsq_cor_ndia_create(int flag, const tjsq_opset* sys_op, const tjsq_opset* env_ops, int env_trans,
                         const tjsq_opset* env_site, int env_site_trans, const tjsq_opset* sys_site, int sys_site_trans,
                         const tjsq_basis* sys_basis, const tjsq_basis* env_basis, int nw, const int* wsec,
                         tjsq_program** prog);

/* The general product of two operators of one side, as the pairing
 * measurements form them (Get_sys_two and Get_sys_two1, code/tjsq/measure.f90:
 * 3798-3889; Get_SCOP_Operator_Env, :6211-6263): each operator of `ops`
 * (block length L - 1) embedded into `basis` (update_block_ndia_*), the site
 * operator `site` embedded (update_site_ndia_*), their coupling to total
 * 2S = `rank` (block_mul_block_ndia with new_oper%down_dif = rank,
 * code/tjsq/block.f90:1956-2197), projected with `trun` (update_trun_dia for
 * a particle-conserving scalar, update_trun_ndia otherwise,
 * code/tjsq/block.f90:1311-1503).  One operator at length L per operator of
 * `ops`.  Run with ops = trun data, x = the data of `ops`. */
int tjsq_product_ndia_create(int side, int flag, int rank, const tjsq_opset* ops, const tjsq_opset* site,
                             const tjsq_basis* basis, const tjsq_opset* trun, tjsq_program** prog);

/* The bond-site contractions of Get_SC_Order (code/tjsq/measure.f90:
 * 3226-3376): kind 0 sys_site_env_site_cor_ndia (:5709-5924) of `sys_site`
 * and `env_site`; kind 1 sys_site_env_block_cor_ndia (:5926-6154) of
 * `sys_site` with every operator of `env_ops` (environment block length
 * env_len - 1); kind 2 env_block_site_cor_ndia (:5085-5238) of every
 * operator of `env_ops` with `env_site`.  *_trans != 0 applies
 * block_transfer_trans (code/tjsq/block.f90:1676-1706) to that operand
 * first.  The output is one vector of the wavefunction's length per
 * operator of env_ops (one for kind 0): each routine's value = psi . y_q.
 * Run with ops = the data of env_ops (any one-element array for kind 0),
 * x = psi, y cleared first. */
int tjsq_site_cor_create(int kind, int flag, const tjsq_opset* sys_site, int sys_site_trans, const tjsq_opset* env_ops,
                         int env_trans, const tjsq_opset* env_site, int env_site_trans, const tjsq_basis* sys_basis,
                         const tjsq_basis* env_basis, int nw, const int* wsec, tjsq_program** prog);

/* Get_density_matrix_sys (side 0) / _env (side 1), levels >= 1 states summed
 * and normalised first (Get_density_matrix_*_new, code/tjsq/den_mat.f90:
 * 4-134).  The output operator has one sub-block per wavefunction sector. */
int tjsq_density_create(int side, int len, int nw, const int* wsec, int levels, tjsq_program** prog);

/* wave_trans_LR (dir 0) / wave_trans_RL (dir 1) (code/tjsq/dmrg_sub.f90:
 * 1093-1238, 1241-1384): the ground state of one bond carried to the next
 * bond of a sweep, the start vector of the next solve (the "initial guess"
 * step before lanczos_memo_new).  sys_bs / env_bs are the fused bases the
 * reference routine takes (LR: new system, old environment; RL: old system,
 * new environment); sys_trun / env_trun hold one operator each;
 * wsec_old / wsec_new are the sector lists (6 ints each) of the old and the
 * new wavefunction.  Run with ops = sys_trun data followed by env_trun data,
 * x = the old wavefunction, y = the new one (normalised, wave_normalize). */
int tjsq_wave_transform_create(int dir, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, const tjsq_opset* sys_trun,
                               const tjsq_opset* env_trun, int nw_old, const int* wsec_old, int nw_new,
                               const int* wsec_new, tjsq_program** prog);

int tjsq_program_sizes(const tjsq_program* prog, int64_t* sizes);
int tjsq_program_layout(const tjsq_program* prog, int* hdr, int* sub, int64_t* off);
int tjsq_program_tables(const tjsq_program* prog, void* seg1, void* tile1, void* seg2, void* tile2, void* rtile);
int tjsq_program_stats(const tjsq_program* prog, tjsq_plan_stats* stats);
int tjsq_program_run(tjsq_program* prog, const double* ops, const double* x, double* y, void* stream);
int tjsq_program_profile(tjsq_program* prog, int enable);
int tjsq_program_profile_read(tjsq_program* prog, double* ms, int64_t* launches);
int tjsq_program_destroy(tjsq_program* prog);
/* tjsq_program_run on host buffers (ops, x in; y out), for callers whose data
 * stay on the host, as the reference's own routines keep theirs: the Fortran
 * drop-in of Get_density_matrix_*_new, Get_Hamiltonian_Sys/Env and
 * wave_trans_LR/RL (code/tjsq/den_mat.f90:84-134, get_sys.f90:4-478,
 * dmrg_sub.f90:1093-1384).  x holds what tjsq_program_run's x holds (for a
 * density program, `levels` states back to back).  The device copy of y
 * starts cleared (a measurement program's y needs it).  Runs on the null
 * stream; returns when y is on the host. */
int tjsq_program_run_host(tjsq_program* prog, const double* ops, const double* x, double* y);

typedef struct tjsq_trunc_info {
  int num_state;       /* kept states (Num_state) */
  int dim;             /* states of the density matrix */
  double trunc_err;    /* discarded weight of the normalised spectrum */
  double entropy;      /* Entropys: -sum w log w of the normalised spectrum */
  double entropy1;     /* entropy with multiplet weights (den_mat.f90:211-222) */
  double entropy2;     /* (den_mat.f90:226-235) */
} tjsq_trunc_info;

/* Get_truncation_operator (code/tjsq/den_mat.f90:141-383) on a density
 * matrix in device memory (den: sub table den_sub[5*nsec], offsets den_off).
 * Each sector is diagonalised in place (ascending eigenvalues, eigenvectors
 * overwrite the sector); the kept states follow the reference: global
 * descending sort (QuickSort_Double order), kept_min, growth while the
 * discarded weight exceeds trun_err, cap kept_max.  The truncation operator
 * is returned as a layout into den: trun_sub[5*trun_num] and trun_off (the
 * kept columns of a sector are its last eigenvectors).  spectrum (host, may
 * be NULL) receives the sorted normalised spectrum. */
int tjsq_truncation_operator(int nsec, const int* den_sub, const int64_t* den_off, double* den, int kept_min,
                             int kept_max, double trun_err, int* trun_sub, int64_t* trun_off, int* trun_num,
                             double* spectrum, tjsq_trunc_info* info, void* stream);

/* tjsq_truncation_operator with den in host memory (the Fortran drop-in of
 * Get_truncation_operator, code/tjsq/den_mat.f90:141-383): den is copied to
 * the device, diagonalised there, and the eigenvectors are copied back over
 * it; trun_off indexes into the host den. */
int tjsq_truncation_operator_host(int nsec, const int* den_sub, const int64_t* den_off, double* den, int kept_min,
                                  int kept_max, double trun_err, int* trun_sub, int64_t* trun_off, int* trun_num,
                                  double* spectrum, tjsq_trunc_info* info);

/* Host part of Get_truncation_operator on given per-sector ascending
 * eigenvalues (no GPU): kept count per sector, sorted spectrum and the
 * 1-based state index of each sorted value (QuickSort_Double's permutation,
 * code/tjsq/block.f90:288-332); any output may be NULL. */
int tjsq_truncation_select(int nsec, const int* den_sub, const double* eigval, int kept_min, int kept_max,
                           double trun_err, int* counts, double* spectrum, int* order, tjsq_trunc_info* info);

/* Identity truncation of the warm-up when the fused basis fits in kept_max
 * (block_transfer_basis + unit matrices, code/tjsq/dmrg_sub.f90:659-667 and
 * code/tjsq/block.f90:1645-1673): layout of one sub-block per basis sector;
 * data (device, may be NULL) receives the unit matrices. */
int tjsq_identity_truncation(const tjsq_basis* basis, int* trun_sub, int64_t* trun_off, int64_t* ndata, double* data,
                             void* stream);

/* ---- batches: many applies in one grouped launch per stage -------------
 * A batch runs the applies of n plans (any sizes and shapes) as one launch of
 * each stage: the members' task tables are merged, their operator arenas
 * copied into one.  x and y hold the members' vectors back to back at the
 * offsets tjsq_batch_offsets returns (n + 1 entries, the last = total).
 * Each member's result is bit-identical to its plan's own apply.  With
 * tjsq_batch_graph(b, 1) the two stages are captured once into a hipGraph
 * (per precision and x / y pair) and replayed.  precision_mask (bits as in
 * tjsq_plan_create) selects the operand copies the batch holds.  Plans may be
 * destroyed after the batch is created. */
typedef struct tjsq_batch tjsq_batch;
int tjsq_batch_create(tjsq_plan* const* plans, int n, int precision_mask, void* stream, tjsq_batch** batch);
int tjsq_batch_offsets(const tjsq_batch* batch, int64_t* off);
int tjsq_batch_stats(const tjsq_batch* batch, tjsq_plan_stats* stats);
int tjsq_batch_graph(tjsq_batch* batch, int enable);
int tjsq_batch_apply(tjsq_batch* batch, const double* x, double* y, void* stream);
int tjsq_batch_apply_f32(tjsq_batch* batch, int precision, const float* x, float* y, void* stream);
int tjsq_batch_profile(tjsq_batch* batch, int enable);
int tjsq_batch_profile_read(tjsq_batch* batch, double* ms, int64_t* launches);
int tjsq_batch_destroy(tjsq_batch* batch);

/* Releases the library's per-device state that outlives single calls: the
 * eigensolver's rocBLAS handles, streams, events and work buffers of
 * tjsq_truncation_operator.  No counterpart in the reference, whose process
 * exit frees everything; a host program calls it before its GPU runtime is
 * torn down (the Python binding registers it at interpreter exit).  Later
 * calls that need the state create it again. */
int tjsq_shutdown(void);

/* Last error message of the calling thread (static storage). */
const char* tjsq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
