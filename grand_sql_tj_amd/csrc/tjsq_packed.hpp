// tjsq_packed.hpp -- packed-operand task tables of the BF16x3 H_eff apply.
//
// The mixed-precision apply reads every operand as "lines": a matrix op(M)
// used along a tile dimension u and a contraction index k is stored as one
// line per u, each line holding k in groups of eight, eight bf16 high halves
// (16 bytes) followed by the eight low halves (16 bytes) of the split
// v = hi + lo.  Every sub-block is zero-padded to multiples of 16 in both
// directions, so a 16 x 16 x 32 MFMA operand of a lane is two aligned
// 16-byte loads and the tile loop needs no bounds tests.
//
// Operators are packed once per plan from the FP64 arena, x once per apply
// from the FP32 vector (both orientations where referenced), and stage 1
// writes its products T = op(A) op(X) directly as lines of T (rows of T,
// contraction over its columns), the operand stage 2 reads.  Stage 1
// computes T^T = op(X)^T op(A)^T so that each lane's accumulator registers
// hold four consecutive k of one line of T.
//
// The contributions, their coefficients and the stage-1 sharing are those
// of compile_program (tjsq_plan.cpp); only the storage of the operands and
// the tiling (one 64-lane wave per output tile of up to 64 x 64) differ.
#pragma once

#include <cstdint>
#include <vector>

#include "tjsq_plan.hpp"

namespace tjsq {

enum : uint8_t { PK_OPS = 0, PK_X = 1, PK_T = 2, PK_XF = 3 };   // PK_XF: the FP32 x itself (axpy)
enum : int32_t { PK_OUT_Y = 0, PK_OUT_T = 1 };

// one packed copy of a sub-block: src (R x C, column-major, leading dim
// src_ld) into `lines` lines of `kpad` elements; trans = 0: line = column of
// the source (k runs down the rows), trans = 1: line = row (k along columns)
struct PackJob {
  int64_t src_off;
  int64_t dst_off;       // bytes
  int32_t src_ld, rows, cols, trans, lines, kpad;
};

// (32-bit fields only: the tile kernel reads segments with scalar loads).
// On the host l_off / r_off are byte offsets into the spaces; the device
// copy holds absolute addresses and, in l_lim / r_lim, the bytes from there
// to the end of the space (PK_XF: l_off = element offset into x, l_stride
// its leading dimension, in both copies)
struct PSeg {
  int64_t l_off, r_off;
  uint32_t l_lim, r_lim;
  int32_t l_stride, r_stride;   // bytes per line
  int32_t K;             // contraction length, a multiple of 16 (0 for axpy)
  int32_t spaces;        // l_space | r_space << 8
  double c;
  int l_space() const { return spaces & 0xff; }
  int r_space() const { return (spaces >> 8) & 0xff; }
};

struct PTile {
  int64_t out_off;       // PK_OUT_Y: element offset of the output block in y; PK_OUT_T: byte offset of T's lines
  int32_t out_ld;        // PK_OUT_Y: leading dimension; PK_OUT_T: bytes per line of T
  int32_t m0, n0;        // tile origin in the (padded) output block
  int32_t mb, nb;        // 16 x 16 blocks of the tile (1..4 each)
  int32_t mlim, nlim;    // rows / columns of the block that exist (relative to the block origin)
  int32_t seg_begin, seg_end;
  int32_t out_kind;
  int32_t slot;          // >= 0: write partial slot `slot` (4096 floats, 64 x 64 column-major)
  int32_t pad0;
};

struct PReduce {         // ordered sum of partial slots [s0, s1) into y
  int64_t out_off;
  int32_t out_ld, m0, n0, mt, nt, s0, s1, pad0;
};

struct PackedPlan {
  std::vector<PackJob> ops_jobs, x_jobs;
  int64_t ops_bytes = 0, x_bytes = 0, t_bytes = 0;
  std::vector<PSeg> seg1, seg2;
  std::vector<PTile> tile1, tile2;
  std::vector<PReduce> rtile;
  int64_t nslots = 0;
  double flops_exec = 0;   // MFMA flops of the padded tiles (one product per algorithmic multiply-add)
  // stage-1 tiles in quads: tiles 4 q .. 4 q + 3 read the same x lines
  // (tile 4 q leads; padding tiles have no blocks) and run as one workgroup
  bool quad1 = false;
};

struct PackShape {
  int k_split = 256;       // K pieces of stage-2 segments (a multiple of 32)
  // K chunks of 32 per stage-2 work item: at the m = 8000 bench bond, 16 /
  // 32 / 64 / 128 gave stage 2 + partial sums of 1.07 / 0.93 / 0.90 / 0.91 ms
  int item_chunks = 64;
  // stage-1 tiles sorted by their x block (sys sector outer, so the
  // operators of one sys sector are read while the sector's x blocks are
  // swept) and then by operator copy; 0: in compile order
  int order1 = 1;
  // with order1: the tiles of one operator copy (the right operand) kept on
  // one XCD, spread over `xcds1` XCDs by work as for stage 2 below.  At the
  // m = 8000 bench bond 8 XCDs took stage 1 from 1.093 to 1.053 ms; grouping
  // by x block instead left it at 1.10 ms, by wave sector or by system
  // sector (fewer, larger groups) made it 1.19 ms.
  int xcds1 = 8;
  // stage-1 tiles that read the same x lines (the same rows of the same x
  // blocks, segment by segment) grouped in fours, one workgroup each, which
  // shares each x chunk through LDS (with xcds1: the quads of one operator
  // copy of their lead tile kept on one XCD)
  int quad1 = 1;
  // stage-2 tiles that read the same left operand first (a T block, or an
  // operator or x block) kept on one XCD, groups spread over `xcds2` XCDs by
  // work (workgroup g runs on XCD g mod xcds2); 1: table order.  At the
  // m = 8000 bench bond 8 XCDs took stage 2 from 0.834 to 0.72 ms (grouping
  // by output column strip, the blocks that share right operators: 0.772 ms).
  int xcds2 = 8;
};

// Builds the packed tables from P.contrib / P.out_blocks (filled by
// build_host_plan).  Returns TJSQ_OK, or TJSQ_ERR_ARG with msg for a
// contribution kind this layout does not cover (the caller then keeps the
// unpacked path).
int compile_packed(const HostPlan& P, const PackShape& shape, PackedPlan& out, std::string& msg);

}  // namespace tjsq

static_assert(sizeof(tjsq::PackJob) == 40, "PackJob is copied to the device verbatim");
static_assert(sizeof(tjsq::PSeg) == 48, "PSeg is copied to the device verbatim");
static_assert(sizeof(tjsq::PTile) == 56, "PTile is copied to the device verbatim");
static_assert(sizeof(tjsq::PReduce) == 40, "PReduce is copied to the device verbatim");
