// tjsq_update.hpp -- density matrix, truncation and block update programs.
//
// After each Lanczos solve the reference (code/tjsq/dmrg_sub.f90:656-693)
// forms the reduced density matrix of the enlarged block
// (Get_density_matrix_sys/env, den_mat.f90:4-134), diagonalises it per
// sector and keeps the largest eigenvectors (Get_truncation_operator,
// den_mat.f90:141-383), then enlarges the block by one site and projects
// every operator with the truncation matrix T (Get_Hamiltonian_Sys/Env,
// get_sys.f90:4-478, with update_block_*, update_site_*, Get_hamiltonian_dia
// and _ndia, update_trun_dia/ndia of block.f90:382-1503).
//
// Here the enlarged operators are never formed.  Every sub-block of an
// enlarged operator is a sum of placements: c * (a sub-block of the old
// block operator, possibly transposed) or c * identity, put at a rectangle
// of the fused basis.  Projecting a placement with T is one GEMM chain
//
//     Y[eff sub-block] += c * T_lhs[rows]^T * src * T_rhs[cols]
//
// (or c * T_lhs[rows]^T * T_rhs[cols] for an identity), which the grouped
// tile engine of tjsq_plan.hpp executes: space SP_OPS holds T, SP_X the old
// block's operator data and the output is the new block's operator data.
#pragma once

#include <string>
#include <vector>

#include "tjsq_plan.hpp"

namespace tjsq {

// Flat layout of a program's output operator set (tjsq_opset layout).
struct OutLayout {
  int nops = 0;
  std::vector<int> hdr;       // 4 per operator
  std::vector<int> sub;       // 5 per sub-block
  std::vector<int64_t> off;   // element offset per sub-block
  int64_t ndata = 0;
};

struct Program {
  HostPlan hp;                // contributions and compiled task tables
  OutLayout out;
  int64_t x_elems = 0;        // elements the X operand must hold
  int64_t ops_elems = 0;      // elements the SP_OPS operand must hold
  int levels = 1;             // density programs: states summed before use
  bool normalize_out = false; // wave transforms: the output is normalised
};

// wave_trans_LR (dir 0, code/tjsq/dmrg_sub.f90:1093-1238) / wave_trans_RL
// (dir 1, :1241-1384): the ground state of one bond carried to the next bond
// of a sweep, the start vector of the next solve.  sys_bs / env_bs are the
// fused bases the reference passes (LR: the new system basis and the old
// environment basis; RL: the old system basis and the new environment
// basis), sys_trun / env_trun one operator each, wsec_old / wsec_new the
// sector lists (6 ints each, Get_super_basis order) of the two wavefunctions.
// The program's operator space is sys_trun's data followed by env_trun's,
// its input the old wavefunction, its output the new one, normalised.
int build_wave_transform_program(int dir, const tjsq_basis& sys_bs, const tjsq_basis& env_bs,
                                 const tjsq_opset& sys_trun, const tjsq_opset& env_trun, int nw_old,
                                 const int* wsec_old, int nw_new, const int* wsec_new, const Shape& shape,
                                 Program& pg, std::string& msg);

// Get_Hamiltonian_Sys (side 0) / Get_Hamiltonian_Env (side 1): pre is the
// block model at length L (nops = 1 + 4 * active sites), basis the fused
// basis of that block with one site (Get_basis), trun the truncation
// operator over the basis sectors (one operator).  The output is the model
// at length L + 1 with nleg11(L + 1) active sites.
int build_update_program(const tjsq_lattice& lat, int side, int sys_len, int env_len, const tjsq_opset& pre,
                         const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                         std::string& msg);

// update_trun_dia (ndia[i] = 0) / update_trun_ndia (ndia[i] = 1) applied to
// every operator of `ops` with the same truncation operator.
int build_projection_program(const tjsq_opset& ops, const int* ndia, const tjsq_opset& trun, const Shape& shape,
                             Program& pg, std::string& msg);

// One step of the transport of diagonal operators to a longer block
// (Get_sys_oper_dia / Get_env_oper_dia, code/tjsq/get_sys.f90:764-891):
// each operator of ops (diagonal, at block length L; may be NULL) embedded
// into `basis` (update_block_dia), then the site operator `site` (one
// operator of 1x1 sub-blocks with host data; may be NULL) embedded
// (update_site_dia), each projected with `trun` (update_trun_dia).  Run
// with ops = trun data, x = the data of ops.
int build_dia_transport_program(const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis& basis,
                                const tjsq_opset& trun, const Shape& shape, Program& pg, std::string& msg);
// The same with the products of the first nprod operators of ops with the
// site operator appended (each embedded as a reduced operator of the fused
// basis, then projected): the two-point operators n_i n_j of a block formed
// when site j joins it.  Output order for nprod > 0: ops[0, nprod), the
// site operator, ops[nprod, end), the products.
int build_dia_transport_program(const tjsq_opset* ops, const tjsq_opset* site, int nprod, const tjsq_basis& basis,
                                const tjsq_opset& trun, const Shape& shape, Program& pg, std::string& msg);

// Expectation values of diagonal block operators and of a site operator in
// a wavefunction (measure_sys_block / _site, measure_env_block / _site,
// code/tjsq/measure.f90:1241-1470): the output is nops + 1 vectors of the
// wavefunction's length, Y_j for every operator of ops and Y_site last,
// with <O_j> = psi . Y_j.  basis is the fused basis of the measured side.
// Run with ops = the operators' data, x = psi, y cleared beforehand.
int build_measure_program(int side, const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis& basis,
                          int nw, const int* wsec, const Shape& shape, Program& pg, std::string& msg);

// One step of the transport of rank-k operators (Get_sys_oper_ndia /
// Get_env_oper_ndia / Change_*_oper_ndia, code/tjsq/get_sys.f90:998-1238):
// each operator of ops (block length L; may be NULL) embedded into `basis`
// (update_block_ndia_sys / _env, side 0 / 1), then the site operator `site`
// (1x1 sub-blocks, host data; may be NULL) embedded (update_site_ndia_*),
// each projected with `trun` (update_trun_ndia).  flag: 'B' or 'F' (the
// fermion signs of the embeddings).  Run with ops = trun data, x = ops data.
int build_ndia_transport_program(int side, char flag, const tjsq_opset* ops, const tjsq_opset* site,
                                 const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                                 std::string& msg);

// Correlations of one system block operator A (block length sys_len - 1)
// with environment block operators (sys_block_env_block_cor_ndia,
// code/tjsq/measure.f90:5241-5480), the environment bond-site operator
// (sys_block_env_site_cor_ndia, :5482-5706) and the system bond-site
// operator (sys_block_site_cor_ndia, :4939-5083): output copy q holds Y_q,
// the routine's value = psi . Y_q (environment block operators first, then
// env_site, then sys_site).  sys_basis / env_basis are the wavefunction's
// fused bases; *_trans applies block_transfer_trans to that operand first.
// Run with ops = A's data followed by the environment operators' data,
// x = psi, y cleared beforehand.
int build_cor_ndia_program(const tjsq_opset& sys_op, const tjsq_opset* env_ops, int env_trans,
                           const tjsq_opset* env_site, int env_site_trans, const tjsq_opset* sys_site,
                           int sys_site_trans, const tjsq_basis& sys_basis, const tjsq_basis& env_basis, int nw,
                           const int* wsec, char flag, const Shape& shape, Program& pg, std::string& msg);

// The same-side pair operators of Get_sys2b / Get_env2b (side 0 / 1,
// code/tjsq/measure.f90:2748-2819) formed when the second site joins the
// block: each operator of ops (rank k, block length L - 1) embedded into
// `basis` (update_block_ndia_*), the site operator (1x1 sub-blocks, host
// data) embedded (update_site_ndia_*), their rank-0 coupling
// (block_mul_block_ndia, code/tjsq/block.f90:1956-2197) projected with
// `trun` (update_trun_dia): one diagonal operator at length L per operator
// of ops.  Run with ops = trun data, x = ops data.
int build_pair_ndia_program(int side, char flag, const tjsq_opset& ops, const tjsq_opset& site,
                            const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                            std::string& msg);

// The general product of two operators of one side (Get_sys_two,
// Get_sys_two1, Get_SCOP_Operator_Env, code/tjsq/measure.f90:3798-3889,
// 6211-6263): each operator of ops (block length L - 1) embedded into
// `basis`, the site operator embedded, their rank-`rank` coupling
// (block_mul_block_ndia, code/tjsq/block.f90:1956-2197), projected with
// `trun` (update_trun_dia for a particle-conserving scalar, update_trun_ndia
// otherwise).  Run with ops = trun data, x = ops data.
int build_product_ndia_program(int side, char flag, int rank, const tjsq_opset& ops, const tjsq_opset& site,
                               const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                               std::string& msg);

// The bond-site contractions of Get_SC_Order (code/tjsq/measure.f90:
// 3226-3376): kind 0 sys_site_env_site_cor_ndia (:5709-5924), kind 1
// sys_site_env_block_cor_ndia (:5926-6154) with every operator of env_ops,
// kind 2 env_block_site_cor_ndia (:5085-5238) with every operator of
// env_ops.  Output copy q holds Y_q with the routine's value = psi . Y_q.
// Run with ops = env_ops' data (kinds 1, 2), x = psi, y cleared first.
int build_site_cor_program(int kind, const tjsq_opset* sys_site, int sys_site_trans, const tjsq_opset* env_ops,
                           int env_trans, const tjsq_opset* env_site, int env_site_trans,
                           const tjsq_basis& sys_basis, const tjsq_basis& env_basis, int nw, const int* wsec,
                           char flag, const Shape& shape, Program& pg, std::string& msg);

// Get_density_matrix_sys (side 0) / _env (side 1) of a wavefunction with
// sectors wsec; `len` is the block length recorded in the output operator.
int build_density_program(int side, int len, int nw, const int* wsec, const Shape& shape, Program& pg,
                          std::string& msg);

// Host part of Get_truncation_operator: eigval holds the eigenvalues of every
// density-matrix sector in ascending order (DSYEV order), sectors back to
// back.  Returns the kept count per sector and the reference's diagnostics.
struct TruncSelection {
  int num_state = 0;
  int dim = 0;
  double trunc_err = 0, entropy = 0, entropy1 = 0, entropy2 = 0;
  std::vector<int> counts;      // kept states per density sector
  std::vector<double> sorted;   // normalised spectrum, descending (QuickSort_Double order)
  std::vector<int> order;       // 1-based global state index of each sorted value
};
int select_truncation(int nsec, const int* den_sub, const double* eigval, int kept_min, int kept_max,
                      double trun_err, TruncSelection& sel, std::string& msg);

// QuickSort_Double (block.f90:288-332): descending Hoare partition with the
// middle element as pivot, on 1-based inclusive ranges of a and idx.
void quicksort_double(double* a, int* idx, int lower, int upper);

}  // namespace tjsq
