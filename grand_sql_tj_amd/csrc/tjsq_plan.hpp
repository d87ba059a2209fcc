// tjsq_plan.hpp -- host-side plan builder for the H_eff apply.
//
// The reference applies H_eff (model_wave, code/tjsq/get_sys.f90:4127-4240)
// push-style: 14 kernels walk the wavefunction sectors, search target sectors
// linearly and call DGEMM per sub-block pair.  Here the same walk runs once
// per superblock on the host and records every contribution
//
//     Y[out block] += c * op(A) * X[in block] * op(B)        (A, B optional)
//
// as a Contrib.  The contributions are then compiled into two task tables
// for the device: stage 1 forms the distinct products T = op(A) X of the
// two-sided contributions, stage 2 pulls, for every output tile, the sum of
// its segments c * op(L) * op(R).  Pull order makes the apply free of atomics
// and deterministic.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "tjsq_capi.h"

namespace tjsq {

// ---- block-sparse views (host metadata; data lives in an arena) ----------
struct OpSub {
  int num_up, num_down, down_dif, row_dim, sdim;
  int64_t off;            // element offset in the operator arena
};
struct Op {
  int len = 0, up_dif = 0, down_dif = 0;
  std::vector<OpSub> sub;
};
struct OpSetView {        // one block model (Total_Model)
  Op ham;
  std::vector<Op> site;   // 4 per active site: elec_up, elec_down, spin_sd, num_sn
  const Op& op(int active_site_1based, int which) const { return site[4 * (active_site_1based - 1) + which]; }
};
struct MapSub { int bl_up, bl_down, st_up, st_down, spos, sdim; };
struct MapSec { int new_up, new_down, dim; std::vector<MapSub> sub; };
struct BasisView { std::vector<MapSec> sec; };
struct WSec { int sys_up, sys_down, env_up, env_down, sys_dim, env_dim; int64_t off; };

// ---- lattice tables (dmrg_setup0 / dmrg_setup) -----------------------------
struct LatticeTables {
  int nsite = 0;
  std::vector<int> neib;      // (nsite+1) x 12
  std::vector<int> neiba;     // (nsite+1)^2
  std::vector<int> nposi, nposi_lat, nposi_dm;
  std::vector<int> inb1;      // (nsite+1)^2, [i*(nsite+1)+ns]
  std::vector<int> nleg11;
  double jt[13], jd[13], jn[13], esite;
  int neib_at(int i, int k) const { return neib[(size_t)i * 12 + k - 1]; }
  int npos(int j) const { return nposi[nposi_dm[j]]; }
  int inb1_at(int i, int ns) const { return inb1[(size_t)i * (nsite + 1) + ns]; }
};
void build_lattice(const tjsq_lattice& lat, LatticeTables& L);
void lattice_setup(LatticeTables& L, int nleg, int ns, int ns21);

struct Couplings {
  int coup_bs[3][3] = {{0}};
  int coup_bb[3] = {0};
  std::vector<int> inb;
};
void hmerge_tables(const LatticeTables& L, int sys_len, int env_len, Couplings& C);

double w6js(int j1, int j2, int j3, int l1, int l2, int l3);
double w3js(int j1, int j2, int j3, int m1, int m2, int m3);

// ---- merged operators (hmerge's snc / sub_s) ------------------------------
// A merged operator copies the structure of active site 1's operator and is
// the coefficient sum of other active sites' operators.  Recipes are
// executed on the device: dst[e] = sum over terms of coef * src[e], terms in
// the reference's accumulation order.
struct MergeTerm { int64_t dst_off, src_off, n; double coef; };

// ---- contributions ---------------------------------------------------------
enum : uint8_t { SP_OPS = 0, SP_X = 1, SP_T = 2, SP_IDENT = 3 };
// C_AXPY:  Y += c X                 C_LEFT:  Y += c op(A) X
// C_RIGHT: Y += c X op(B)           C_BOTH:  Y += c op(A) op(X) op(B)
// C_PROD:  Y += c op(A) op(B), A and B read from spaces a_space, b_space
enum : uint8_t { C_AXPY = 0, C_LEFT = 1, C_RIGHT = 2, C_BOTH = 3, C_PROD = 4 };

struct Contrib {
  uint8_t kind;
  uint8_t a_t, b_t;         // op(A) = A^T if a_t, op(B) = B^T if b_t
  uint8_t x_t;              // op(X) = X^T (C_BOTH)
  uint8_t a_space, b_space; // operand spaces of C_PROD (SP_OPS or SP_X)
  int64_t a_off, b_off;     // operator arena offsets
  int a_ld, b_ld;           // leading dimensions of the stored A, B
  int64_t x_off; int x_ld;  // input sub-block, op(X) is K1 x K2
  int out_block;            // output sub-block id
  int M, N, K1, K2;
  double c;
};

struct OutBlock { int64_t off; int ld, M, N, group; };  // group: (wave sector, env sub-block)

// device task tables (POD, copied verbatim)
struct Seg {
  int64_t l_off, r_off;
  int32_t l_ld, r_ld, K;
  uint8_t l_space, r_space, l_t, r_t;
  double c;
};
struct Tile {
  int64_t out_off;
  int32_t out_ld, m0, n0, mt, nt, seg_begin, seg_end;
  int32_t out_space;
  int32_t slot;             // >= 0: write a partial tile to partial slot `slot`
  int32_t group;            // locality group (work items sharing operands), informational
};

struct HostPlan {
  // inputs
  int sys_len = 0, env_len = 0;
  OpSetView sys, env;
  BasisView sys_bs, env_bs;
  std::vector<WSec> wsec;
  int64_t dim = 0;
  int64_t arena_elems = 0;      // sys data + env data + merged
  int64_t sys_base = 0, env_base = 0, merged_base = 0;
  // outputs
  std::vector<MergeTerm> merge;
  std::vector<OutBlock> out_blocks;
  std::vector<Contrib> contrib;
  std::vector<Seg> seg1, seg2;
  std::vector<Tile> tile1, tile2;
  // stage-2 work items that share an output tile write partial tiles; rtile
  // sums slots [seg_begin, seg_end) of each such tile in order
  std::vector<Tile> rtile;
  int64_t nslots = 0;           // partial storage in units of 32x32 elements
  int n64_1 = 0, n64_2 = 0;     // leading 64x64 work items of tile1 / tile2
  int64_t scratch_elems = 0;
  tjsq_plan_stats stats{};
  std::vector<int64_t> ops_used; // distinct operator sub-blocks referenced (elements)
};

// Builds the whole host plan.  Returns TJSQ_OK or an error code with msg set.
// Work-item shaping of stage 2: segments longer than k_split are cut into K
// pieces and consecutive pieces of an output tile are grouped into work items
// of at most `item_chunks` K-chunks of `kc` (an axpy counts one chunk).
struct Shape {
  int big = 64;             // tile edge for blocks with a side above `small`
  int small = 32;
  int kc = 32;
  int k_split = 256;
  int item_chunks = 32;
  int xcds = 1;             // > 1: items of a locality group go to one XCD (round-robin dispatch);
                            // the plain order measured faster on MI355X, so it is the default
  int merge_right = 1;      // factor two-sided terms of one output block that share op(B)
                            // (6-8 % faster applies at the m = 8000 workload)
  int order = 0;            // stage-2 item order: 0 = output blocks in wave-sector order;
                            // 1 = column strips (blocks sharing an env sub-block) together,
                            // each block's segments sorted by their right operand, so the
                            // strip's items walk the shared env operators in one order
};

// Compiles P.contrib and P.out_blocks into seg1/tile1/seg2/tile2/rtile and
// the stats' flop and operand counts.  Every output block is written exactly
// once (a block without contributions is written with zeros).
int compile_program(HostPlan& P, const Shape& shape);

int build_host_plan(const tjsq_lattice& lat, int sys_len, int env_len, const tjsq_opset& sys,
                    const tjsq_opset& env, const tjsq_basis& sys_bs, const tjsq_basis& env_bs, int nw,
                    const int* wsec, const Shape& shape, HostPlan& P, std::string& msg);

int get_basis(int nb, const int* blk, int nsite, const int* site, std::vector<int>& sec,
              std::vector<int>& sub);
int get_super_basis(const tjsq_basis& s, const tjsq_basis& e, int total_up, int total_down,
                    std::vector<int>& wsec);

}  // namespace tjsq

static_assert(sizeof(tjsq::Seg) == 40, "Seg layout is part of the C-ABI");
static_assert(sizeof(tjsq::Tile) == 48, "Tile layout is part of the C-ABI");
static_assert(sizeof(tjsq::MergeTerm) == 32, "MergeTerm layout is part of the C-ABI");
