// tjsq_hip.hip -- CDNA4 (gfx950) kernels and the C-ABI of libtjsq_hip.so.
//
// The H_eff apply runs as two grouped-GEMM launches over the task tables of
// tjsq_plan.cpp:
//   stage 1  T_j = op(A_j) X_j            for the distinct A.X products of the
//                                         two-sided terms (block x block);
//   stage 2  Y_tile = sum_s c_s op(L_s) op(R_s)  pulled per output tile,
//                                         no atomics, fixed summation order.
// Each workgroup is one 64-lane wavefront that owns one 32x32 output tile and
// accumulates it in MFMA registers.  Result layouts (measured on gfx950): lane
// l, register r of a 16x16 block holds row l/16 + 4r for the FP64 16x16x4
// instruction and row 4(l/16) + r for the FP32 and BF16 ones, column l%16 in
// all three; operand A row i / k-group g comes from lane i + 16g (v_mfma_f64_16x16x4f64 in FP64,
// v_mfma_f32_16x16x4f32 in FP32, v_mfma_f32_16x16x16bf16_1k three times for
// the BF16x3 split), staging 32x16 / 16x32 operand panels through LDS.
//
// The Lanczos solver (lanczos_memo_new, code/tjsq/eigen.f90:2-175) keeps its
// Krylov basis on the device; dot products are two-pass deterministic
// reductions in FP64.
#include <hip/hip_runtime.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "tjsq_capi.h"
#include "tjsq_plan.hpp"
#include "tjsq_update.hpp"
#include "tjsq_packed.hpp"

using tjsq::HostPlan;
using tjsq::Seg;
using tjsq::Tile;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(TJSQ_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)


typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// operand words and spaces
// ---------------------------------------------------------------------------
enum { MODE_F64 = 0, MODE_F32 = 1, MODE_BF16X3 = 2 };

// T: vectors, accumulators and outputs.  W: one operand element as the tile
// kernel reads it; in BF16x3 a 32-bit word with bf16(v) in the high half and
// bf16(v - bf16(v)) in the low half, split once when the operand is made
// (operators at plan creation, x per apply, T in stage 1's epilogue), so the
// tile loop does no conversion work.
template <int MODE>
struct KT {
  using T = typename std::conditional<MODE == MODE_F64, double, float>::type;
  using W = typename std::conditional<MODE == MODE_F64, double,
                                      typename std::conditional<MODE == MODE_F32, float, uint32_t>::type>::type;
  using Acc = typename std::conditional<MODE == MODE_F64, d4, f4>::type;
};

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }
// v - bf16(v) is exact in FP32, so hi + lo carries v to 2^-16 relative
__device__ __forceinline__ uint32_t split_word(float v) {
  const unsigned short h = bf16_rne(v);
  return ((uint32_t)h << 16) | bf16_rne(v - bf16_to_f(h));
}

// the spaces a launch reads and writes: operators and x as operand words,
// x itself (axpy terms add it directly), stage-1 scratch (operand words of
// T), and the output
template <int MODE>
struct Spaces {
  const typename KT<MODE>::W* ops;
  const typename KT<MODE>::W* xw;
  const typename KT<MODE>::T* x;
  typename KT<MODE>::W* t;
  typename KT<MODE>::T* y;
};

template <int MODE>
__device__ __forceinline__ const typename KT<MODE>::W* space_ptr(const Spaces<MODE>& S, int sp) {
  return sp == tjsq::SP_OPS ? S.ops : (sp == tjsq::SP_X ? S.xw : (const typename KT<MODE>::W*)S.t);
}

// row of register r of lane (lq, lr) inside a 16x16 MFMA result block
template <int MODE>
__device__ __forceinline__ int acc_row(int lq, int r) {
  return MODE == MODE_F64 ? lq + 4 * r : lq * 4 + r;
}

// ---------------------------------------------------------------------------
// Grouped tile kernel (one template for FP64, FP32 and BF16x3)
//
// A workgroup owns one TSZ x TSZ output tile (TSZ = 64: four waves, one 32x32
// quarter each; TSZ = 32: one wave) and walks its segment list in K chunks of
// KC.  Operand elements come in through raw buffer loads: a lane whose
// element lies outside the tile or past the segment's K gets an offset above
// the record count and reads zero, so there are no divergent branches
// around the loads.  Each lane's offsets are one base plus wave-uniform
// steps, set up once per segment.  Chunk c+1 is loaded into registers while
// the MFMAs of chunk c run from LDS.  Both panels sit in LDS as [tile index][k]
// rows; an MFMA operand is 16 (FP32 / BF16x3) or 32 (FP64) consecutive bytes
// of a row per lane.  The k order inside an MFMA step is a permutation shared
// by both operands, so the sum over k is unchanged.
// ---------------------------------------------------------------------------
template <int MODE, int TSZ>
struct TileCfg {
  static constexpr int NT = TSZ == 64 ? 256 : 64;    // threads
  // K chunk staged per step: 32 (FP32 / BF16x3), 16 (FP64, half the
  // prefetch registers: 5 waves per SIMD instead of 4, 6 % faster).  A chunk
  // of 64 for FP32 / BF16x3 cost one wave per SIMD and measured slower.
  static constexpr int KC = MODE == MODE_F64 ? 16 : 32;
  static constexpr int PER = TSZ * KC / NT;          // elements per thread per panel and chunk
  static_assert(KC % (MODE == MODE_BF16X3 ? 32 : 16) == 0 && PER >= 1, "K chunk must be a multiple of the MFMA k group");
  // LDS row stride in operand words: rows start 16-byte aligned and the
  // operand reads of 8 consecutive lanes fall in distinct banks
  static constexpr int HS = MODE == MODE_F64 ? KC + 2 : KC + 4;
  // waves per SIMD asked of the compiler for the 64-tiles (the kernel is
  // latency-bound): 5 (BF16x3 87 VGPRs; 6 spills and runs slower; FP64
  // 96 VGPRs with 5 spilled, still faster than 4 waves without spills)
  static constexpr int WAVES = TSZ == 64 ? 5 : 4;
};

constexpr uint32_t OOB = 0x80000000u;   // at or above the record count: reads as zero

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7FFFFFFF, 0x00020000);
}

template <typename W>
__device__ __forceinline__ W buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(W) == 8)
    return __builtin_bit_cast(W, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  else
    return __builtin_bit_cast(W, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// One operand panel of a segment as the threads load it: TSZ tile rows (or
// columns) by KC values of k per chunk.  A k-major operand (contiguous along
// k) gives each lane a fixed k = tid % KC and tile index tid / KC + q NT / KC.
// A tile-major one (contiguous along the tile) gives the lane tile index
// a = tid % TSZ and k = tid / TSZ + q NT / TSZ for 8-byte words; for 4-byte
// words a = tid % 16 + 16 (q % G) and k = (tid / 16) % 4 + 4 wave +
// 4 NW (q / G), G = TSZ / 16, NW = waves, so that one store instruction
// writes 16 rows x 4 k, which fall in distinct LDS banks (a store of 64 rows
// at one k would hit only 16 banks with the 16-byte aligned row stride).
template <int MODE>
constexpr bool quad_map() { return MODE != MODE_F64; }

template <int MODE, int TSZ>
struct Side {
  __amdgpu_buffer_rsrc_t r;
  uint32_t lane;   // byte offset of the lane's first element at k0 = 0
  uint32_t qs;     // byte step between the lane's elements
  uint32_t ks;     // byte step per K chunk
  int fix;         // the lane's tile index (tile-major) or k (k-major)
  int lim;         // tile extent
  bool kmaj;
};

template <int MODE, int TSZ>
__device__ __forceinline__ Side<MODE, TSZ> make_side(const typename KT<MODE>::W* base, int64_t off, int ld,
                                                     bool kmaj, int t0, int lim, int tid) {
  using C = TileCfg<MODE, TSZ>;
  constexpr uint32_t WB = sizeof(typename KT<MODE>::W);
  Side<MODE, TSZ> s;
  s.r = make_rsrc(base + off);
  s.lim = lim;
  s.kmaj = kmaj;
  if (!kmaj && quad_map<MODE>()) {
    const int a = tid % 16, b = (tid / 16) % 4 + 4 * (tid / 64);
    s.fix = a;
    s.lane = (uint32_t)((t0 + a) + b * ld) * WB;
    s.qs = (uint32_t)(4 * (C::NT / 64) * ld) * WB;
    s.ks = (uint32_t)(C::KC * ld) * WB;
  } else if (!kmaj) {
    const int a = tid % TSZ, b = tid / TSZ;
    s.fix = a;
    s.lane = (uint32_t)((t0 + a) + b * ld) * WB;
    s.qs = (uint32_t)((C::NT / TSZ) * ld) * WB;
    s.ks = (uint32_t)(C::KC * ld) * WB;
  } else {
    const int b = tid % C::KC, a = tid / C::KC;
    s.fix = b;
    s.lane = (uint32_t)(b + (t0 + a) * ld) * WB;
    s.qs = (uint32_t)((C::NT / C::KC) * ld) * WB;
    s.ks = (uint32_t)C::KC * WB;
  }
  return s;
}

// the lane's elements of the chunk starting kv values before the segment's
// end (kv = K - k0), kofs = chunk index * ks
template <int MODE, int TSZ>
__device__ __forceinline__ void load_side(const Side<MODE, TSZ>& s, int kv, uint32_t kofs, int tid,
                                          typename KT<MODE>::W* v) {
  using C = TileCfg<MODE, TSZ>;
  using W = typename KT<MODE>::W;
  if (!s.kmaj && quad_map<MODE>()) {
    constexpr int G = TSZ / 16;
    constexpr uint32_t WB = sizeof(W);
    const int b0 = (tid / 16) % 4 + 4 * (tid / 64);
#pragma unroll
    for (int q = 0; q < C::PER; ++q) {
      const bool ok = s.fix + 16 * (q % G) < s.lim && b0 + 4 * (C::NT / 64) * (q / G) < kv;
      v[q] = buf_load<W>(s.r, ok ? s.lane + kofs + (uint32_t)(16 * (q % G)) * WB + (uint32_t)(q / G) * s.qs : OOB);
    }
  } else if (!s.kmaj) {
    // 64-tiles: b0 = tid / 64 is the wave index (scalar)
    const int b0 = TSZ == 64 ? __builtin_amdgcn_readfirstlane(tid / TSZ) : tid / TSZ;
    const bool oka = s.fix < s.lim;
#pragma unroll
    for (int q = 0; q < C::PER; ++q) {
      const bool ok = oka && b0 + q * (C::NT / TSZ) < kv;
      v[q] = buf_load<W>(s.r, ok ? s.lane + kofs + (uint32_t)q * s.qs : OOB);
    }
  } else {
    const int a0 = tid / C::KC;
    const bool okk = s.fix < kv;
#pragma unroll
    for (int q = 0; q < C::PER; ++q) {
      const bool ok = okk && a0 + q * (C::NT / C::KC) < s.lim;
      v[q] = buf_load<W>(s.r, ok ? s.lane + kofs + (uint32_t)q * s.qs : OOB);
    }
  }
}

// the loaded elements into the [tile index][k] LDS panel (FP32 / FP64 left
// operands scaled by the segment coefficient on the way)
template <int MODE, int TSZ, bool SCALE>
__device__ __forceinline__ void store_side(typename KT<MODE>::W* lds, bool kmaj, int tid,
                                           const typename KT<MODE>::W* v, typename KT<MODE>::T c) {
  using C = TileCfg<MODE, TSZ>;
  auto val = [&](int q) {
    if constexpr (SCALE && MODE != MODE_BF16X3) return v[q] * c;
    else return v[q];
  };
  if (!kmaj && quad_map<MODE>()) {
    constexpr int G = TSZ / 16;
    const int a = tid % 16, b = (tid / 16) % 4 + 4 * (tid / 64);
#pragma unroll
    for (int q = 0; q < C::PER; ++q) lds[(a + 16 * (q % G)) * C::HS + b + 4 * (C::NT / 64) * (q / G)] = val(q);
  } else if (!kmaj) {
    const int a = tid % TSZ, b0 = tid / TSZ;
#pragma unroll
    for (int q = 0; q < C::PER; ++q) lds[a * C::HS + b0 + q * (C::NT / TSZ)] = val(q);
  } else {
    const int b = tid % C::KC, a0 = tid / C::KC;
#pragma unroll
    for (int q = 0; q < C::PER; ++q) lds[(a0 + q * (C::NT / C::KC)) * C::HS + b] = val(q);
  }
}

// eight operand words (consecutive k) -> the hi and lo bf16 operands
__device__ __forceinline__ void split_pairs(u4 x0, u4 x1, bf8& hi, bf8& lo) {
  constexpr unsigned HI = 0x07060302u, LO = 0x05040100u;
  u4 h, l;
  h.x = __builtin_amdgcn_perm(x0.y, x0.x, HI); l.x = __builtin_amdgcn_perm(x0.y, x0.x, LO);
  h.y = __builtin_amdgcn_perm(x0.w, x0.z, HI); l.y = __builtin_amdgcn_perm(x0.w, x0.z, LO);
  h.z = __builtin_amdgcn_perm(x1.y, x1.x, HI); l.z = __builtin_amdgcn_perm(x1.y, x1.x, LO);
  h.w = __builtin_amdgcn_perm(x1.w, x1.z, HI); l.w = __builtin_amdgcn_perm(x1.w, x1.z, LO);
  hi = __builtin_bit_cast(bf8, h);
  lo = __builtin_bit_cast(bf8, l);
}

// STAGE (1 or 2) does not change the code: it gives the two stages distinct
// kernel symbols, so a trace reports them apart.
template <int MODE, int TSZ, int STAGE>
__global__ __launch_bounds__((TileCfg<MODE, TSZ>::NT), (TileCfg<MODE, TSZ>::WAVES)) void k_tiles(
    const Tile* __restrict__ tiles, const Seg* __restrict__ segs, Spaces<MODE> S, typename KT<MODE>::T* partial,
    int ntiles, const int* __restrict__ gate) {
  using T = typename KT<MODE>::T;
  using W = typename KT<MODE>::W;
  using Acc = typename KT<MODE>::Acc;
  using C = TileCfg<MODE, TSZ>;
  constexpr int PER = C::PER, KC = C::KC, HS = C::HS;
  const int bid = blockIdx.x;
  if (bid >= ntiles || (gate && *gate)) return;
  const Tile tl = tiles[bid];
  const int tid = threadIdx.x;
  // the wave index is uniform per wave: read it into a scalar register so
  // the tile-bound tests on it are scalar branches
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (TSZ == 64) ? (wave >> 1) * 32 : 0, wn = (TSZ == 64) ? (wave & 1) * 32 : 0;
  const int lr = lane & 15, lq = lane >> 4;
  __shared__ __align__(16) W Ls[TSZ * HS];
  __shared__ __align__(16) W Rs[TSZ * HS];
  Acc acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r) acc[a][b][r] = T(0);

  int si = tl.seg_begin;
  const int se = tl.seg_end;
  for (; si < se; ++si) {  // axpy segments lead every list; they add x itself
    const Seg s = segs[si];
    if (s.r_space != tjsq::SP_IDENT) break;
    const T c = (T)s.c;
    for (int bi = 0; bi < 2; ++bi)
      for (int bj = 0; bj < 2; ++bj)
        for (int r = 0; r < 4; ++r) {
          const int row = wm + bi * 16 + acc_row<MODE>(lq, r), col = wn + bj * 16 + lr;
          if (row < tl.mt && col < tl.nt)
            acc[bi][bj][r] += c * S.x[s.l_off + (tl.m0 + row) + (int64_t)(tl.n0 + col) * s.l_ld];
        }
  }
  if (si < se) {
    // BF16x3 operand words cannot be scaled by the segment coefficient, so
    // the running sum is kept as scale * acc: at a segment's first chunk acc
    // is multiplied by scale / c and scale becomes c (coefficients are never
    // 0: zero terms are not emitted), and the MFMAs then add A.B unscaled;
    // acc *= scale at the end.  One accumulator set instead of two keeps the
    // kernel at 87 VGPRs, 5 waves per SIMD.
    double scale = 1.0;
    Seg cur = segs[si];
    // descriptor of the segment after `cur`, loaded one segment ahead so the
    // chunk loads at a segment boundary do not wait on it
    Seg pre = segs[min(si + 1, se - 1)];
    // (segments pass by value: a referenced descriptor would live in
    // per-lane memory and lose its wave-uniformity)
    auto side_l = [&](Seg g) {
      return make_side<MODE, TSZ>(space_ptr(S, g.l_space), g.l_off, g.l_ld, g.l_t != 0, tl.m0, tl.mt, tid);
    };
    auto side_r = [&](Seg g) {
      return make_side<MODE, TSZ>(space_ptr(S, g.r_space), g.r_off, g.r_ld, g.r_t == 0, tl.n0, tl.nt, tid);
    };
    Side<MODE, TSZ> Lc = side_l(cur), Rc = side_r(cur);
    int k0 = 0;
    uint32_t kc = 0;   // chunk index inside the segment
    W rl[PER], rr[PER];
    load_side(Lc, cur.K, 0u, tid, rl);
    load_side(Rc, cur.K, 0u, tid, rr);
    for (;;) {
      store_side<MODE, TSZ, true>(Ls, Lc.kmaj, tid, rl, (T)cur.c);
      store_side<MODE, TSZ, false>(Rs, Rc.kmaj, tid, rr, T(1));
      __syncthreads();
      // the chunk in LDS is (cur, k0); the loads below fetch the next one
      // (the sides move on to the next segment at a segment's last chunk)
      const int kvalid = min(KC, cur.K - k0);
      const bool first_chunk = k0 == 0;
      const bool last_chunk = k0 + KC >= cur.K;
      const double cseg = cur.c;
      if (last_chunk) {
        ++si;
        k0 = 0;
        kc = 0;
        cur = pre;
        pre = segs[min(si + 1, se - 1)];
        if (si < se) {
          Lc = side_l(cur);
          Rc = side_r(cur);
        }
      } else {
        k0 += KC;
        ++kc;
      }
      const bool more = si < se;
      if (more) {
        load_side(Lc, cur.K - k0, kc * Lc.ks, tid, rl);
        load_side(Rc, cur.K - k0, kc * Rc.ks, tid, rr);
      }
      // 16-row / 16-column blocks outside the tile hold only zeros: skip them
      // (wave-uniform); so are MFMA k groups past the chunk's valid K
      const bool r0 = wm < tl.mt, r1 = wm + 16 < tl.mt, c0 = wn < tl.nt, c1 = wn + 16 < tl.nt;
      const bool rr_[2] = {r0, r1}, cc_[2] = {c0, c1};
      if constexpr (MODE == MODE_BF16X3) {
        if (first_chunk) {
          const float rs = (float)(scale / cseg);
          scale = cseg;
          for (int bi = 0; bi < 2; ++bi)
            for (int bj = 0; bj < 2; ++bj)
              for (int r = 0; r < 4; ++r) acc[bi][bj][r] *= rs;
        }
        // k step g of 32: lane (lr, lq) supplies k = 32 g + 8 lq + t, t < 8
        for (int g = 0; g * 32 < kvalid; ++g) {
          bf8 bh[2], bl[2];
          for (int p = 0; p < 2; ++p) {
            const W* rb = Rs + (wn + p * 16 + lr) * HS + g * 32 + lq * 8;
            split_pairs(*(const u4*)rb, *(const u4*)(rb + 4), bh[p], bl[p]);
          }
          for (int bi = 0; bi < 2; ++bi) {
            if (!rr_[bi]) continue;
            bf8 ah, al;
            const W* la = Ls + (wm + bi * 16 + lr) * HS + g * 32 + lq * 8;
            split_pairs(*(const u4*)la, *(const u4*)(la + 4), ah, al);
            for (int bj = 0; bj < 2; ++bj) {
              if (!cc_[bj]) continue;
              acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[bj], acc[bi][bj], 0, 0, 0);
              acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[bj], acc[bi][bj], 0, 0, 0);
              acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[bj], acc[bi][bj], 0, 0, 0);
            }
          }
        }
      } else {
        // k group g of 16: lane (lr, lq) supplies k = 16 g + 4 lq + s in step s
        // (FP64 in two halves of two steps, one 16-byte read per row each)
        constexpr int SH = MODE == MODE_F64 ? 2 : 4;
        for (int g = 0; g * 16 < kvalid; ++g)
          for (int h = 0; h < 4; h += SH) {
            T a[2][SH], b[2][SH];
            for (int p = 0; p < 2; ++p) {
              const W* la = Ls + (wm + p * 16 + lr) * HS + g * 16 + lq * 4 + h;
              const W* rb = Rs + (wn + p * 16 + lr) * HS + g * 16 + lq * 4 + h;
              if constexpr (MODE == MODE_F64) {
                const d2 av = *(const d2*)la, bv = *(const d2*)rb;
                a[p][0] = av.x; a[p][1] = av.y;
                b[p][0] = bv.x; b[p][1] = bv.y;
              } else {
                const f4 av = *(const f4*)la, bv = *(const f4*)rb;
                for (int q = 0; q < 4; ++q) { a[p][q] = av[q]; b[p][q] = bv[q]; }
              }
            }
            for (int q = 0; q < SH; ++q)
              for (int bi = 0; bi < 2; ++bi)
                for (int bj = 0; bj < 2; ++bj) {
                  if (!(rr_[bi] && cc_[bj])) continue;
                  if constexpr (MODE == MODE_F64)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi][q], b[bj][q], acc[bi][bj], 0, 0, 0);
                  else
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[bi][q], b[bj][q], acc[bi][bj], 0, 0, 0);
                }
          }
      }
      __syncthreads();
      if (!more) break;
    }
    if constexpr (MODE == MODE_BF16X3) {
      const float rs = (float)scale;
      for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj)
          for (int r = 0; r < 4; ++r) acc[bi][bj][r] *= rs;
    }
  }
  if (tl.slot >= 0) {            // partial tile: slot counts units of 32x32 elements
    const int64_t base = (int64_t)tl.slot * 1024;
    for (int bi = 0; bi < 2; ++bi)
      for (int bj = 0; bj < 2; ++bj)
        for (int r = 0; r < 4; ++r) {
          const int row = wm + bi * 16 + acc_row<MODE>(lq, r), col = wn + bj * 16 + lr;
          if (row < tl.mt && col < tl.nt) partial[base + row + (int64_t)col * TSZ] = acc[bi][bj][r];
        }
    return;
  }
  const int64_t base = tl.out_off + tl.m0 + (int64_t)tl.n0 * tl.out_ld;
  const int ld = tl.out_ld;
  const bool to_t = tl.out_space == tjsq::SP_T;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int r = 0; r < 4; ++r) {
        const int row = wm + bi * 16 + acc_row<MODE>(lq, r), col = wn + bj * 16 + lr;
        if (!(row < tl.mt && col < tl.nt)) continue;
        const int64_t e = base + row + (int64_t)col * ld;
        if (to_t) {               // stage-1 product: stored as operand words
          if constexpr (MODE == MODE_BF16X3) S.t[e] = split_word(acc[bi][bj][r]);
          else S.t[e] = acc[bi][bj][r];
        } else {
          S.y[e] = acc[bi][bj][r];
        }
      }
}

// ordered sum of the partial tiles of split work items; a reduce tile lists
// its partial tiles as slot ranges [seg_begin, seg_end) of `slot` units each
// (slot = 1 for 32x32 partials, 4 for 64x64) with leading dimension out_space
template <typename T>
__global__ __launch_bounds__(256) void k_reduce(const Tile* __restrict__ rt, const T* __restrict__ partial, T* y,
                                                int n, const int* __restrict__ gate) {
  const int t = blockIdx.x;
  if (t >= n || (gate && *gate)) return;
  const Tile tl = rt[t];
  const int units = tl.slot, pld = tl.out_space;
  const int np = (tl.seg_end - tl.seg_begin) / units;
  const int64_t step = (int64_t)units * 1024;
  for (int e = threadIdx.x; e < tl.mt * tl.nt; e += blockDim.x) {
    const int row = e % tl.mt, col = e / tl.mt;
    const T* p = partial + (int64_t)tl.seg_begin * 1024 + row + col * pld;
    // four independent loads in flight, added in ascending partial order
    T acc = T(0);
    int q = 0;
    for (; q + 4 <= np; q += 4) {
      const T a0 = p[q * step], a1 = p[(q + 1) * step], a2 = p[(q + 2) * step], a3 = p[(q + 3) * step];
      acc += a0;
      acc += a1;
      acc += a2;
      acc += a3;
    }
    for (; q < np; ++q) acc += p[q * step];
    y[tl.out_off + (tl.m0 + row) + (int64_t)(tl.n0 + col) * tl.out_ld] = acc;
  }
}

// ---------------------------------------------------------------------------
// Packed BF16x3 apply (tables of tjsq_packed.cpp)
//
// Operands are lines of k in groups of eight: 16 bytes of bf16 high halves,
// then 16 bytes of low halves.  One 64-lane wave owns one output tile of up
// to 4 x 4 blocks of 16 x 16 and loads each MFMA operand straight into its
// registers: lane (i = lane % 16, q = lane / 16) reads line i of a block,
// group q of the 32-wide K chunk, as two aligned 16-byte buffer loads.
// Everything is zero-padded to multiples of 16, so the loop has no bounds
// tests; a K chunk that is only 16 long (the padded K is a multiple of 16)
// zeroes the left operand of lanes q >= 2 (the right one reads finite data:
// the next line or zeros past the buffer end).  The next chunk's operands,
// also across segments, are loaded while the current chunk's MFMAs run.
// ---------------------------------------------------------------------------
// the vectors of one packed launch: x itself (axpy segments) and y; the
// operand copies and the stage-1 products are addressed by the tables
struct PSpaces {
  const float* xf;
  float* y;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc_n(const void* base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nbytes, 0x00020000);
}

// A packing job is cut into pieces of PACK_PIECE (line, group) items, one
// piece per workgroup (grid-stride over the pieces), so a large block is
// spread over many CUs instead of serialising one workgroup.
constexpr int PACK_PIECE = 256;
struct PackPiece { int job, e0; };

// a thread packs one group of eight k of one line (device jobs hold
// absolute destinations).  Consecutive threads take consecutive source
// elements: along a line where k runs down the source rows (trans = 0),
// across lines where it runs along them (trans = 1).
template <typename S>
__global__ __launch_bounds__(PACK_PIECE) void k_pack(const tjsq::PackJob* __restrict__ jobs,
                                                     const PackPiece* __restrict__ pieces, int npieces,
                                                     const S* __restrict__ src, const int* __restrict__ gate) {
  if (gate && *gate) return;
  for (int pc = blockIdx.x; pc < npieces; pc += gridDim.x) {
    const PackPiece P = pieces[pc];
    const tjsq::PackJob J = jobs[P.job];
    const int groups = J.kpad / 8;
    const int n = J.lines * groups;
    const int e = P.e0 + threadIdx.x;
    if (e < n) {
      const int l = J.trans ? e % J.lines : e / groups, g = J.trans ? e / J.lines : e % groups;
      unsigned short h[8], w[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int k = 8 * g + t;
        const int r = J.trans ? l : k, c = J.trans ? k : l;
        const bool in = r < J.rows && c < J.cols;
        if constexpr (sizeof(S) == 8) {
          // FP64 operators keep their FP64 residual in the low half (as the
          // unpacked operand words do)
          const double v = in ? (double)src[J.src_off + r + (int64_t)c * J.src_ld] : 0.0;
          h[t] = bf16_rne((float)v);
          w[t] = bf16_rne((float)(v - (double)bf16_to_f(h[t])));
        } else {
          const float v = in ? (float)src[J.src_off + r + (int64_t)c * J.src_ld] : 0.0f;
          h[t] = bf16_rne(v);
          w[t] = bf16_rne(v - bf16_to_f(h[t]));   // exact in FP32
        }
      }
      u4 hi, lo;
      hi.x = h[0] | ((uint32_t)h[1] << 16); hi.y = h[2] | ((uint32_t)h[3] << 16);
      hi.z = h[4] | ((uint32_t)h[5] << 16); hi.w = h[6] | ((uint32_t)h[7] << 16);
      lo.x = w[0] | ((uint32_t)w[1] << 16); lo.y = w[2] | ((uint32_t)w[3] << 16);
      lo.z = w[4] | ((uint32_t)w[5] << 16); lo.w = w[6] | ((uint32_t)w[7] << 16);
      u4* d = (u4*)((uint8_t*)J.dst_off + (int64_t)l * J.kpad * 4 + (int64_t)g * 32);   // absolute address
      d[0] = hi;
      d[1] = lo;
    }
  }
}

std::vector<PackPiece> pack_pieces(const std::vector<tjsq::PackJob>& jobs) {
  std::vector<PackPiece> out;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const int n = jobs[j].lines * (jobs[j].kpad / 8);
    for (int e0 = 0; e0 < n; e0 += PACK_PIECE) out.push_back({(int)j, e0});
  }
  return out;
}

struct PSide {
  __amdgpu_buffer_rsrc_t r;
  uint32_t v[4];
};

__device__ __forceinline__ PSide packed_side(int64_t addr, uint32_t lim, int stride, int origin, int nblk, int li,
                                             int lq) {
  PSide s;
  s.r = make_rsrc_n((const void*)addr, lim);
  const uint32_t v0 = (uint32_t)(origin + li) * (uint32_t)stride + (uint32_t)lq * 32u;
#pragma unroll
  for (int b = 0; b < 4; ++b) s.v[b] = b < nblk ? v0 + (uint32_t)(16 * b) * (uint32_t)stride : OOB;
  return s;
}

// high and low halves of the lane's eight k of block b, ko bytes into the
// lines (all offsets in the lane's VGPR, where the range check sees them)
__device__ __forceinline__ void ld_op(const PSide& s, int b, uint32_t ko, u4& hi, u4& lo) {
  hi = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(s.r, (int)(s.v[b] + ko), 0, 0));
  lo = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(s.r, (int)(s.v[b] + ko + 16u), 0, 0));
}

// stage 1's output: lane (li, lq) holds D[u][v] = T[v][u] for u = m0 + 16 bi
// + 4 lq + r, v = n0 + 16 bj + li: four consecutive k of line v of T, in
// group m0 / 8 + 2 bi + lq / 2, half lq % 2 of its eight, written as the
// next stage's packed operand (bf16 high and low halves)
__device__ __forceinline__ void store_t_lines(const tjsq::PTile& tl, const f4 (&acc)[4][4], int mb, int nb, int li,
                                              int lq) {
  uint8_t* tb = (uint8_t*)tl.out_off + (int64_t)(tl.n0 + li) * tl.out_ld + (tl.m0 / 8 + lq / 2) * 32 + (lq % 2) * 8;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) {
      if (bi >= mb || bj >= nb) continue;
      unsigned short h[4], w[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        h[r] = bf16_rne(acc[bi][bj][r]);
        w[r] = bf16_rne(acc[bi][bj][r] - bf16_to_f(h[r]));
      }
      uint8_t* d = tb + (int64_t)(16 * bj) * tl.out_ld + 64 * bi;
      *(uint2*)d = uint2{h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16)};
      *(uint2*)(d + 16) = uint2{w[0] | ((uint32_t)w[1] << 16), w[2] | ((uint32_t)w[3] << 16)};
    }
}

// one output tile of a packed stage (64 lanes)
template <int STAGE>
__device__ __forceinline__ void packed_tile(const tjsq::PTile& tl, const tjsq::PSeg* __restrict__ segs, const PSpaces& S,
                                            float* __restrict__ partial) {
  const int lane = threadIdx.x, li = lane & 15, lq = lane >> 4;
  const int mb = tl.mb, nb = tl.nb;
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  // the lane's rows (4 consecutive, + 16 bi) and column (+ 16 bj) in the tile
  const int row0 = tl.m0 + 4 * lq, col0 = tl.n0 + li;
  int si = tl.seg_begin;
  const int se = tl.seg_end;
  for (; si < se; ++si) {   // axpy segments lead the list; they add x itself
    const tjsq::PSeg s = segs[si];
    if ((s.spaces & 0xff) != tjsq::PK_XF) break;
    const float c = (float)s.c;
    const float* xb = S.xf + s.l_off;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        if (bi >= mb || bj >= nb) continue;
        const int col = col0 + 16 * bj;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + 16 * bi + r;
          if (row < tl.mlim && col < tl.nlim) acc[bi][bj][r] += c * xb[row + (int64_t)col * s.l_stride];
        }
      }
  }
  double scale = 1.0;
  if (si < se) {
    tjsq::PSeg cur = segs[si];
    tjsq::PSeg nxt = segs[min(si + 1, se - 1)];
    PSide L = packed_side(cur.l_off, cur.l_lim, cur.l_stride, tl.m0, mb, li, lq);
    PSide R = packed_side(cur.r_off, cur.r_lim, cur.r_stride, tl.n0, nb, li, lq);
    int k0 = 0;
    // two register sets, A and B: one chunk is multiplied while the next
    // one loads into the other set
    u4 aAh[4], aAl[4], bAh[4], bAl[4], aBh[4], aBl[4], bBh[4], bBl[4];
    auto issue = [&](uint32_t ko, u4 (&ah)[4], u4 (&al)[4], u4 (&bh)[4], u4 (&bl)[4]) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        ld_op(L, b, ko, ah[b], al[b]);
        ld_op(R, b, ko, bh[b], bl[b]);
      }
    };
    // moves the sides past the chunk (cur, k0): false after the last one,
    // when the sides read nothing (every block out of range)
    auto advance = [&](uint32_t& ko) -> bool {
      int kn = k0 + 32;
      bool more = true;
      if (kn >= cur.K) {
        ++si;
        kn = 0;
        more = si < se;
        cur = nxt;
        nxt = segs[min(si + 1, se - 1)];
        L = packed_side(cur.l_off, cur.l_lim, cur.l_stride, tl.m0, more ? mb : 0, li, lq);
        R = packed_side(cur.r_off, cur.r_lim, cur.r_stride, tl.n0, more ? nb : 0, li, lq);
      }
      k0 = kn;
      ko = (uint32_t)kn * 4u;
      return more;
    };
    auto compute = [&](bool first, bool half, double cseg, u4 (&ah)[4], u4 (&al)[4], const u4 (&bh)[4],
                       const u4 (&bl)[4]) {
      if (first) {
        // the running sum is kept as scale * acc (operand words cannot be
        // scaled): acc *= scale / c, scale = c; coefficients are never 0
        const float rs = (float)(scale / cseg);
        scale = cseg;
        if (rs != 1.0f) {
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] *= rs;
        }
      }
      if (half && lq >= 2) {   // a 16-long chunk: lanes supplying k 16..31 add nothing
#pragma unroll
        for (int b = 0; b < 4; ++b) { ah[b] = u4{0u, 0u, 0u, 0u}; al[b] = u4{0u, 0u, 0u, 0u}; }
      }
#pragma unroll
      for (int bi = 0; bi < 4; ++bi) {
        if (bi >= mb) continue;
        const bf8 xh = __builtin_bit_cast(bf8, ah[bi]), xl = __builtin_bit_cast(bf8, al[bi]);
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
          if (bj >= nb) continue;
          const bf8 yh = __builtin_bit_cast(bf8, bh[bj]), yl = __builtin_bit_cast(bf8, bl[bj]);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, yh, acc[bi][bj], 0, 0, 0);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yl, acc[bi][bj], 0, 0, 0);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, yh, acc[bi][bj], 0, 0, 0);
        }
      }
    };
    issue(0u, aAh, aAl, bAh, bAl);
    for (;;) {
      bool first = k0 == 0, half = cur.K - k0 == 16;
      double cseg = cur.c;
      uint32_t ko;
      bool more = advance(ko);
      issue(ko, aBh, aBl, bBh, bBl);
      compute(first, half, cseg, aAh, aAl, bAh, bAl);
      if (!more) break;
      first = k0 == 0;
      half = cur.K - k0 == 16;
      cseg = cur.c;
      more = advance(ko);
      issue(ko, aAh, aAl, bAh, bAl);
      compute(first, half, cseg, aBh, aBl, bBh, bBl);
      if (!more) break;
    }
    const float rs = (float)scale;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] *= rs;
  }
  if (tl.out_kind == tjsq::PK_OUT_T) {
    store_t_lines(tl, acc, mb, nb, li, lq);
    return;
  }
  if (tl.slot >= 0) {   // partial tile: 64 x 64, column-major, tile coordinates
    float* pb = partial + (int64_t)tl.slot * 4096 + 4 * lq + li * 64;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        if (bi >= mb || bj >= nb) continue;
        *(f4*)(pb + 16 * bi + 16 * 64 * bj) = acc[bi][bj];
      }
    return;
  }
  float* yb = S.y + tl.out_off + (int64_t)col0 * tl.out_ld + row0;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) {
      if (bi >= mb || bj >= nb) continue;
      if (col0 + 16 * bj >= tl.nlim) continue;
      float* d = yb + (int64_t)(16 * bj) * tl.out_ld + 16 * bi;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (row0 + 16 * bi + r < tl.mlim) d[r] = acc[bi][bj][r];
    }
}

// one workgroup per tile
template <int STAGE>
__global__ __launch_bounds__(64, 2) void k_packed(const tjsq::PTile* __restrict__ tiles, const tjsq::PSeg* __restrict__ segs,
                                                 PSpaces S, float* __restrict__ partial, int ntiles,
                                                 const int* __restrict__ gate) {
  const int bid = blockIdx.x;
  if (bid >= ntiles || (gate && *gate)) return;
  const tjsq::PTile tl = tiles[bid];
  packed_tile<STAGE>(tl, segs, S, partial);
}

// Stage 1 by groups of G tiles.  The G waves of a workgroup form G tiles of
// T that read the same x lines, segment by segment (the same rows of the
// same x blocks, the same K), each with its own operator copy and
// coefficients.  Wave w loads the x blocks b with b % G == w of each chunk
// and stores them to LDS; after one barrier per chunk every wave reads the
// chunk's x blocks from LDS, so the group fetches each x byte from L2 once
// instead of G times.  The chunk sequence is the lead tile's (tile G q), the
// same for all, so every wave reaches every barrier; a padding wave (mb = 0)
// loads its share, synchronises and forms nothing.  LDS holds two chunks
// (written while the other is read), lines padded to 144 bytes so that the
// 16-byte reads and writes of a wave meet no bank twice.
constexpr int QL_LINE = 144;
constexpr int QL_BUF = 4 * 16 * QL_LINE;

template <int G>
__global__ __launch_bounds__(64 * G, 2) void k_packed_group(const tjsq::PTile* __restrict__ tiles,
                                                               const tjsq::PSeg* __restrict__ segs, int ngroups,
                                                               const int* __restrict__ gate) {
  __shared__ __align__(16) uint8_t lds[2 * QL_BUF];
  const int q = blockIdx.x;
  if (q >= ngroups || (gate && *gate)) return;   // the whole workgroup or none of it
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
  const tjsq::PTile lead = tiles[G * q];
  const tjsq::PTile tl = tiles[G * q + w];
  const bool mine = tl.mb > 0;
  const int mb = lead.mb, nb = mine ? tl.nb : 0;
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  const int g0 = lead.seg_begin, gm = tl.seg_begin;
  const int ns = lead.seg_end - g0;
  // the current segment: its x side (this wave's blocks) and its operator side
  int j = 0, k0 = 0;
  tjsq::PSeg ls{};   // (no segment: nothing is loaded)
  if (ns > 0) ls = segs[g0];
  tjsq::PSeg rs = mine && ns > 0 ? segs[gm] : ls;
  __amdgpu_buffer_rsrc_t xr = make_rsrc_n((const void*)ls.l_off, ls.l_lim);
  uint32_t xv0 = (uint32_t)(lead.m0 + li) * (uint32_t)ls.l_stride + (uint32_t)lq * 32u;
  PSide R = packed_side(rs.r_off, rs.r_lim, rs.r_stride, tl.n0, nb, li, lq);
  constexpr int NS = 4 / G;   // x blocks per wave: w, w + G, ...
  // two register sets, A and B: the next chunk loads into one while the
  // current one is multiplied from the other
  u4 xhA[NS], xlA[NS], xhB[NS], xlB[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) xhA[t] = xlA[t] = xhB[t] = xlB[t] = u4{0u, 0u, 0u, 0u};
  u4 rhA[4], rlA[4], rhB[4], rlB[4];
  auto issue = [&](uint32_t ko, u4 (&xh)[NS], u4 (&xl)[NS], u4 (&rh)[4], u4 (&rl)[4]) {
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      const int b = w + G * t;
      if (b < mb) {
        const uint32_t v = xv0 + (uint32_t)(16 * b) * (uint32_t)ls.l_stride + ko;
        xh[t] = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(xr, (int)v, 0, 0));
        xl[t] = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(xr, (int)(v + 16u), 0, 0));
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) ld_op(R, b, ko, rh[b], rl[b]);
  };
  int nc = 0;
  for (int t = 0; t < ns; ++t) nc += (segs[g0 + t].K + 31) / 32;
  double scale = 1.0;
  // chunk c: this wave's x blocks to LDS, the next chunk's loads into the
  // other set, one barrier, then the chunk's products
  auto step = [&](int c, const u4 (&xh)[NS], const u4 (&xl)[NS], const u4 (&rh)[4], const u4 (&rl)[4],
                  u4 (&nxh)[NS], u4 (&nxl)[NS], u4 (&nrh)[4], u4 (&nrl)[4]) {
    const bool first = k0 == 0, half = ls.K - k0 == 16;
    const double cseg = rs.c;
    uint8_t* buf = lds + (c & 1) * QL_BUF;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      const int b = w + G * t;
      if (b < mb) {
        u4* d = (u4*)(buf + (b * 16 + li) * QL_LINE + lq * 32);
        d[0] = xh[t];
        d[1] = xl[t];
      }
    }
    if (c + 1 < nc) {   // the next chunk: past a segment's end, the next segment's sides
      k0 += 32;
      if (k0 >= ls.K) {
        ++j;
        k0 = 0;
        ls = segs[g0 + j];
        rs = mine ? segs[gm + j] : ls;
        xr = make_rsrc_n((const void*)ls.l_off, ls.l_lim);
        xv0 = (uint32_t)(lead.m0 + li) * (uint32_t)ls.l_stride + (uint32_t)lq * 32u;
        R = packed_side(rs.r_off, rs.r_lim, rs.r_stride, tl.n0, nb, li, lq);
      }
      issue((uint32_t)k0 * 4u, nxh, nxl, nrh, nrl);
    }
    __syncthreads();
    u4 ah[4], al[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      ah[b] = u4{0u, 0u, 0u, 0u};
      al[b] = u4{0u, 0u, 0u, 0u};
      if (b < mb && !(half && lq >= 2)) {   // a 16-long chunk: lanes supplying k 16..31 add nothing
        const u4* s4 = (const u4*)(buf + (b * 16 + li) * QL_LINE + lq * 32);
        ah[b] = s4[0];
        al[b] = s4[1];
      }
    }
    if (!mine) return;
    if (first) {
      // the running sum is kept as scale * acc: acc *= scale / c, scale = c
      const float r = (float)(scale / cseg);
      scale = cseg;
      if (r != 1.0f) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] *= r;
      }
    }
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      if (bi >= mb) continue;
      const bf8 xh8 = __builtin_bit_cast(bf8, ah[bi]), xl8 = __builtin_bit_cast(bf8, al[bi]);
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        if (bj >= nb) continue;
        const bf8 yh = __builtin_bit_cast(bf8, rh[bj]), yl = __builtin_bit_cast(bf8, rl[bj]);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl8, yh, acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh8, yl, acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh8, yh, acc[bi][bj], 0, 0, 0);
      }
    }
  };
  if (nc > 0) issue(0u, xhA, xlA, rhA, rlA);
  for (int c = 0; c < nc; c += 2) {
    step(c, xhA, xlA, rhA, rlA, xhB, xlB, rhB, rlB);
    if (c + 1 >= nc) break;
    step(c + 1, xhB, xlB, rhB, rlB, xhA, xlA, rhA, rlA);
  }
  if (!mine) return;
  if (nc > 0) {
    const float r = (float)scale;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] *= r;
  }
  store_t_lines(tl, acc, mb, nb, li, lq);
}

// ordered sum of the partial slots of split packed tiles: thread t sums
// rows 16 (t % 4) .. + 15 of column t / 4 of the 64 x 64 slots, four slots'
// loads in flight at a time, each element's adds in slot order
__global__ __launch_bounds__(256) void k_preduce(const tjsq::PReduce* __restrict__ rt, const float* __restrict__ partial,
                                                 float* __restrict__ y, int n, const int* __restrict__ gate) {
  const int t = blockIdx.x;
  if (t >= n || (gate && *gate)) return;
  const tjsq::PReduce R = rt[t];
  const int col = threadIdx.x >> 2, r0 = 16 * (threadIdx.x & 3);
  if (col >= R.nt || r0 >= R.mt) return;
  const f4* p = (const f4*)(partial + (int64_t)R.s0 * 4096 + r0 + col * 64);
  const int ns = R.s1 - R.s0;
  f4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
  int s = 0;
  for (; s + 4 <= ns; s += 4) {
    f4 v[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) v[u][q] = p[(int64_t)(s + u) * 1024 + q];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] += v[u][q];
  }
  for (; s < ns; ++s)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += p[(int64_t)s * 1024 + q];
  float* d = y + R.out_off + (R.m0 + r0) + (int64_t)(R.n0 + col) * R.out_ld;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (r0 + 4 * q + e < R.mt) d[4 * q + e] = acc[q][e];
}

// ---------------------------------------------------------------------------
// merged operators: dst[e] = sum_t coef_t * src_t[e] in recipe order
// ---------------------------------------------------------------------------
struct MergeDst { int64_t dst_off, n; int t_begin, t_end; };

__global__ void k_merge(double* arena, const MergeDst* dsts, const tjsq::MergeTerm* terms, int ndst) {
  for (int d = blockIdx.y; d < ndst; d += gridDim.y) {
    const MergeDst D = dsts[d];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < D.n; e += (int64_t)gridDim.x * blockDim.x) {
      double acc = 0.0;
      for (int t = D.t_begin; t < D.t_end; ++t) acc = acc + terms[t].coef * arena[terms[t].src_off + e];
      arena[D.dst_off + e] = acc;
    }
  }
}

__global__ void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
__global__ void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}
// BF16x3 operand words of FP64 operators: the low half is the residual of
// the FP64 value
__global__ void k_f64_to_split(const double* __restrict__ a, uint32_t* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = a[i];
    const unsigned short h = bf16_rne((float)v);
    b[i] = ((uint32_t)h << 16) | bf16_rne((float)(v - (double)bf16_to_f(h)));
  }
}
__global__ void k_f32_to_split(const float* __restrict__ a, uint32_t* __restrict__ b, int64_t n,
                               const int* __restrict__ gate) {
  if (gate && *gate) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = split_word(a[i]);
}

// ---------------------------------------------------------------------------
// Lanczos vector kernels (FP64 arithmetic on FP64 or FP32 storage)
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS = 2048;
constexpr int RED_THREADS = 256;
constexpr int RED_MAXV = 256;     // vectors one k_multidot launch takes (max_iter + 1 <= 251)

template <typename T>
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < RED_THREADS / 64; ++q) t += red[q];
  __syncthreads();
  return t;
}

// partial[b * nv + j] = sum over block b's slice of V_j . w  (j < nv), and,
// when with_self, partial[(RED_BLOCKS + b) * nv + j] = V_j . V_j.  Each block
// owns one contiguous slice of the vectors and walks them four at a time
// (w's slice stays in cache); each wave reduces its sums by shuffles into
// LDS, and after one barrier the block's four wave sums are added in a fixed
// order, so the result does not depend on timing.
template <typename T, typename TW = T>
__global__ __launch_bounds__(RED_THREADS) void k_multidot(const T* __restrict__ V, int64_t ldv, int nv,
                                                          const TW* __restrict__ w, int64_t n, double* partial,
                                                          int with_self) {
  constexpr int G = 4;
  __shared__ double red[2][RED_THREADS / 64][RED_MAXV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per, hi = min(n, lo + per);
  for (int j0 = 0; j0 < nv; j0 += G) {
    double s[G], ss[G];
#pragma unroll
    for (int g = 0; g < G; ++g) s[g] = ss[g] = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += RED_THREADS) {
      const double wi = (double)w[i];
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (j0 + g < nv) {
          const double vi = (double)V[(int64_t)(j0 + g) * ldv + i];
          s[g] += vi * wi;
          ss[g] += vi * vi;
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      double a = s[g], b = ss[g];
      for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
      }
      if (lane == 0 && j0 + g < nv) {
        red[0][wid][j0 + g] = a;
        red[1][wid][j0 + g] = b;
      }
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nv; j += RED_THREADS) {
    partial[(int64_t)blockIdx.x * nv + j] = ((red[0][0][j] + red[0][1][j]) + red[0][2][j]) + red[0][3][j];
    if (with_self)
      partial[((int64_t)RED_BLOCKS + blockIdx.x) * nv + j] = ((red[1][0][j] + red[1][1][j]) + red[1][2][j]) + red[1][3][j];
  }
}

// out[j] = sum_b partial[b * nv + j] (fixed tree order); block j (< nv) or
// nv + j for the self products
__global__ __launch_bounds__(RED_THREADS) void k_finish_dots(const double* partial, int nv, int nblocks, int with_self,
                                                             double* out) {
  __shared__ double red[RED_THREADS / 64];
  const int j = blockIdx.x % nv, pass = blockIdx.x / nv;
  if (pass == 1 && !with_self) return;
  double t = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += RED_THREADS) t += partial[((int64_t)pass * RED_BLOCKS + b) * nv + j];
  const double u = block_sum<double>(t, red);
  if (threadIdx.x == 0) out[pass * nv + j] = u;
}

// w = (w - sum_j coef_j V_j) * post_scale ; coefficients from device memory
template <typename T, typename TV = T>
__global__ void k_combine(T* __restrict__ w, const TV* __restrict__ V, int64_t ldv, int nv, const double* coef,
                          int64_t n, double post_scale, int accumulate_into_zero) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = accumulate_into_zero ? 0.0 : (double)w[i];
    double sub = 0.0;
    for (int j = 0; j < nv; ++j) sub += coef[j] * (double)V[(int64_t)j * ldv + i];
    w[i] = (T)((acc - sub) * post_scale);
  }
}

// w = src - sum_j coef_j V_j in FP64 into an FP64 w (src and V in the
// basis type)
template <typename TV>
__global__ void k_combine_from(double* __restrict__ w, const TV* __restrict__ src, const TV* __restrict__ V,
                               int64_t ldv, int nv, const double* coef, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double sub = 0.0;
    for (int j = 0; j < nv; ++j) sub += coef[j] * (double)V[(int64_t)j * ldv + i];
    w[i] = (double)src[i] - sub;
  }
}

template <typename T>
__global__ void k_scale(T* __restrict__ w, int64_t n, double s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = (T)((double)w[i] * s);
}

// w *= 1 / d[0]: the factor k_scale receives from the host, formed on the device
template <typename T>
__global__ void k_scale_inv_dev(T* __restrict__ w, int64_t n, const double* __restrict__ d) {
  const double s = 1.0 / d[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = (T)((double)w[i] * s);
}

// dst[0] = sqrt(src[0]) (a norm from its squared dot)
__global__ void k_sqrt_to(double* dst, const double* src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = sqrt(src[0]);
}

// Gram-Schmidt coefficients <v_j, w> / <v_j, v_j> from the dot outputs
__global__ void k_gs_coef(double* coef, const double* dots, int nv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nv) coef[j] = dots[j] / dots[nv + j];
}

// ---- one Lanczos iteration in five streaming passes ------------------------
// The flags of a pass (device ints): LZ_DONE set once the pass has converged
// or broken down (every later kernel of the pass, the applies included,
// returns at once), LZ_LAST the number of basis vectors, LZ_BRK breakdown.
constexpr int LZ_DONE = 0, LZ_LAST = 1, LZ_BRK = 2;
// slices of the Gram-Schmidt dot kernel and vectors per block
constexpr int LZ_DOT_BLOCKS = 1024, LZ_DOT_GROUP = 16;
// threads (bisection points per round) of the convergence check, and the
// largest tridiagonal it holds (the basis limit, 250 at most)
constexpr int LZ_CHECK_THREADS = 256, LZ_MAXM = 256;

// sum of this block's values of v (one per thread) in a fixed order, given
// to every thread (red: RED_THREADS / 64 + 1 doubles)
__device__ __forceinline__ double block_sum_fixed(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < RED_THREADS / 64; ++q) t += red[q];
    red[RED_THREADS / 64] = t;
  }
  __syncthreads();
  const double t = red[RED_THREADS / 64];
  __syncthreads();
  return t;
}

// W = y - sum_q c_q V_q (nv = 1 or 2 vectors, c from device memory), in
// FP64; partial[b] = this block's part of |W|^2
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_lz_start(double* __restrict__ W, const T* __restrict__ y,
                                                          const T* __restrict__ V, int64_t ldv, int nv,
                                                          const double* __restrict__ c, int64_t n,
                                                          double* __restrict__ partial, const int* __restrict__ flags) {
  __shared__ double red[RED_THREADS / 64 + 1];
  if (flags[LZ_DONE]) return;
  const double c0 = c[0], c1 = nv > 1 ? c[1] : 0.0;
  double ss = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double sub = c0 * (double)V[i];
    if (nv > 1) sub += c1 * (double)V[ldv + i];
    const double w = (double)y[i] - sub;
    W[i] = w;
    ss += w * w;
  }
  const double t = block_sum_fixed(ss, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// partial[b * nv + j] = slice b's part of V_j . W (j < nv): block (b, q)
// takes the vectors 16q .. 16q + 15 over slice b of the elements, each
// thread one accumulator per vector, so W is read once per 16 vectors and
// every block reduces once
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_lz_dots(const T* __restrict__ V, int64_t ldv, int nv,
                                                         const double* __restrict__ W, int64_t n,
                                                         double* __restrict__ partial, const int* __restrict__ flags) {
  constexpr int G = LZ_DOT_GROUP;
  __shared__ double red[RED_THREADS / 64][G];
  if (flags[LZ_DONE]) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int j0 = blockIdx.y * G;
  const int gn = min(G, nv - j0);
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per, hi = min(n, lo + per);
  double s[G];
#pragma unroll
  for (int g = 0; g < G; ++g) s[g] = 0.0;
  const T* v0 = V + (int64_t)j0 * ldv;
  int64_t i = lo + threadIdx.x;
  for (; i + RED_THREADS < hi; i += 2 * RED_THREADS) {
    const double w0 = W[i], w1 = W[i + RED_THREADS];
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < gn) {
        const T* v = v0 + (int64_t)g * ldv + i;
        s[g] += (double)v[0] * w0;
        s[g] += (double)v[RED_THREADS] * w1;
      }
  }
  if (i < hi) {
    const double w0 = W[i];
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < gn) s[g] += (double)v0[(int64_t)g * ldv + i] * w0;
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    double a = s[g];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if (lane == 0) red[wid][g] = a;
  }
  __syncthreads();
  if (threadIdx.x < gn)
    partial[(int64_t)blockIdx.x * nv + j0 + threadIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// beta = |W| from the partials of k_lz_start (every block forms it the same
// way; block 0 stores it, and a beta at or below tol, tmp[5], ends the pass:
// eigen.f90:74-77) and the Gram-Schmidt coefficients of the normalised
// direction w = W / beta (block.f90:2515-2551):
// coef_j = (V_j . W / beta) / (V_j . V_j), block j
__global__ __launch_bounds__(RED_THREADS) void k_lz_coefs(const double* __restrict__ partial, int nv, int nb,
                                                          const double* __restrict__ wpart, int nbw,
                                                          double* __restrict__ beta, const double* __restrict__ vv,
                                                          double* __restrict__ coef, const double* __restrict__ tmp,
                                                          int* __restrict__ flags) {
  __shared__ double red[RED_THREADS / 64 + 1];
  if (flags[LZ_DONE]) return;
  const int j = blockIdx.x;
  double t = 0.0;
  for (int b = threadIdx.x; b < nbw; b += RED_THREADS) t += wpart[b];
  const double bt = sqrt(block_sum_fixed(t, red));
  t = 0.0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) t += partial[(int64_t)b * nv + j];
  const double u = block_sum_fixed(t, red);
  if (threadIdx.x == 0) {
    coef[j] = (u / bt) / vv[j];
    if (j == 0) {
      beta[0] = bt;
      if (fabs(bt) <= tmp[5]) {
        flags[LZ_BRK] = 1;
        flags[LZ_DONE] = 1;
      }
    }
  }
}

// W = W / beta - sum_j coef_j V_j in FP64; partial[b] = block part of |W|^2
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_lz_subtract(double* __restrict__ W, const T* __restrict__ V,
                                                             int64_t ldv, int nv, const double* __restrict__ coef,
                                                             const double* __restrict__ beta, int64_t n,
                                                             double* __restrict__ partial,
                                                             const int* __restrict__ flags) {
  __shared__ double red[RED_THREADS / 64 + 1];
  __shared__ double cs[RED_MAXV];
  if (flags[LZ_DONE]) return;
  for (int j = threadIdx.x; j < nv; j += RED_THREADS) cs[j] = coef[j];
  __syncthreads();
  const double inv = 1.0 / beta[0];
  double ss = 0.0;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * step) {
    double acc[4], sub[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc[u] = i0 + u * step < n ? W[i0 + u * step] * inv : 0.0;
      sub[u] = 0.0;
    }
    for (int j = 0; j < nv; ++j) {
      const T* v = V + (int64_t)j * ldv + i0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i0 + u * step < n) sub[u] += cs[j] * (double)v[u * step];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u * step < n) {
        const double w = acc[u] - sub[u];
        W[i0 + u * step] = w;
        ss += w * w;
      }
  }
  const double t = block_sum_fixed(ss, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// v = W / |W| stored in the basis type (|W| from the partials of
// k_lz_subtract, formed the same way by every block); partial[b] = block
// part of v . v (of the stored values, the denominator later Gram-Schmidt
// steps use)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_lz_store(const double* __restrict__ W, T* __restrict__ v, int64_t n,
                                                          const double* __restrict__ wpart, int nbw,
                                                          double* __restrict__ partial, const int* __restrict__ flags) {
  __shared__ double red[RED_THREADS / 64 + 1];
  if (flags[LZ_DONE]) return;
  double t0 = 0.0;
  for (int b = threadIdx.x; b < nbw; b += RED_THREADS) t0 += wpart[b];
  const double inv = 1.0 / sqrt(block_sum_fixed(t0, red));
  double ss = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T x = (T)(W[i] * inv);
    v[i] = x;
    ss += (double)x * (double)x;
  }
  const double t = block_sum_fixed(ss, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// out = sum of nb partials (a dot or a self dot of the pass)
__global__ __launch_bounds__(RED_THREADS) void k_lz_sum(const double* __restrict__ partial, int nb,
                                                        double* __restrict__ out, const int* __restrict__ flags) {
  __shared__ double red[RED_THREADS / 64 + 1];
  if (flags[LZ_DONE]) return;
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) t += partial[b];
  const double u = block_sum_fixed(t, red);
  if (threadIdx.x == 0) out[0] = u;
}

// partial[b] = block part of a . b (two basis-type vectors)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_lz_dot2(const T* __restrict__ a, const T* __restrict__ b, int64_t n,
                                                         double* __restrict__ partial, const int* __restrict__ flags) {
  __shared__ double red[RED_THREADS / 64 + 1];
  if (flags[LZ_DONE]) return;
  double ss = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    ss += (double)a[i] * (double)b[i];
  const double t = block_sum_fixed(ss, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Sturm count: eigenvalues below x of the m x m tridiagonal with diagonal
// d[i] and squared off-diagonal e2[i] (between i and i + 1), i = 1..m, held
// in LDS; LDL^T pivots guarded by pivmin (the recurrence of LAPACK's DLAEBZ)
__device__ int lz_sturm(const double* __restrict__ d, const double* __restrict__ e2, int m, double x,
                        double pivmin) {
  int c = 0;
  double q = d[1] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++c;
  for (int i = 2; i <= m; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++c;
  }
  return c;
}

// alpha_m = sum of the partials (<v, H v>) and v . v of the stored vector
// from its own partials, then the lowest eigenvalue of
// the m x m tridiagonal by multisection (LZ_CHECK_THREADS points a round) inside
// [theta_prev - w, theta_prev] (interlacing: the lowest Ritz value only
// falls), widened to the Gershgorin interval when the counts reject it; the
// pass converges when it moved by at most max(tol, rtol |theta|)
// (eigen.f90:93-96) or the basis is full.  tmp[0] holds the previous Ritz
// value, tmp[1] its last change, tmp[5..7] tol, rtol and the basis limit.
__global__ __launch_bounds__(LZ_CHECK_THREADS) void k_lz_check(const double* __restrict__ partial, int nb,
                                                               const double* __restrict__ vpart,
                                                               double* __restrict__ vv_out, double* __restrict__ ab,
                                                               int m, double* __restrict__ tmp,
                                                               int* __restrict__ flags) {
  constexpr int NT = LZ_CHECK_THREADS, NW = NT / 64;
  __shared__ double sh[4];
  __shared__ double red[3][NW];
  __shared__ int first_w[NW];
  __shared__ double d[LZ_MAXM + 1], e2[LZ_MAXM + 1];
  if (flags[LZ_DONE]) return;
  const double tol = tmp[5], rtol = tmp[6];
  const int max_iter = (int)tmp[7];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double t = 0.0, tv = 0.0;
  for (int b = tid; b < nb; b += NT) {
    t += partial[b];
    tv += vpart[b];
  }
  for (int o = 32; o > 0; o >>= 1) {
    t += __shfl_down(t, o, 64);
    tv += __shfl_down(tv, o, 64);
  }
  if (lane == 0) {
    red[0][wid] = t;
    red[1][wid] = tv;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, v = 0.0;
    for (int w = 0; w < NW; ++w) {
      a += red[0][w];
      v += red[1][w];
    }
    vv_out[0] = v;
    ab[2 * m - 1] = a;
    sh[3] = a;
  }
  __syncthreads();
  // the tridiagonal into LDS, and the Gershgorin interval and the pivot
  // guard from it (min / max reductions: any order gives the same values)
  double glo = 1e300, ghi = -1e300, emx = 1.0;
  for (int i = tid + 1; i <= m; i += NT) {
    const double di = i == m ? sh[3] : ab[2 * i - 1];
    const double er = i < m ? ab[2 * i] : 0.0, el = i > 1 ? ab[2 * i - 2] : 0.0;
    d[i] = di;
    e2[i] = er * er;
    glo = fmin(glo, di - fabs(el) - fabs(er));
    ghi = fmax(ghi, di + fabs(el) + fabs(er));
    emx = fmax(emx, er * er);
  }
  for (int o = 32; o > 0; o >>= 1) {
    glo = fmin(glo, __shfl_down(glo, o, 64));
    ghi = fmax(ghi, __shfl_down(ghi, o, 64));
    emx = fmax(emx, __shfl_down(emx, o, 64));
  }
  if (lane == 0) {
    red[0][wid] = glo;
    red[1][wid] = ghi;
    red[2][wid] = emx;
  }
  __syncthreads();
  if (tid == 0) {
    double g0 = red[0][0], g1 = red[1][0], em = red[2][0];
    for (int w = 1; w < NW; ++w) {
      g0 = fmin(g0, red[0][w]);
      g1 = fmax(g1, red[1][w]);
      em = fmax(em, red[2][w]);
    }
    sh[0] = g0;
    sh[1] = g1;
    sh[2] = 2.2250738585072014e-308 * em;   // pivmin
  }
  __syncthreads();
  const double gl = sh[0], gu = sh[1], pivmin = sh[2];
  const double prev = tmp[0], moved = tmp[1];
  const double tiny = 4.0 * 2.220446049250313e-16 * fmax(fabs(gl), fabs(gu)) + pivmin;
  double hi = fmin(gu, prev + tiny), lo = fmax(gl, prev - fmax(4.0 * moved, 1e-9 * (1.0 + fabs(prev))));
  // accept the bracket only if it holds exactly the lowest eigenvalue
  if (lz_sturm(d, e2, m, lo, pivmin) != 0) lo = gl - tiny;
  if (lz_sturm(d, e2, m, hi, pivmin) < 1) hi = gu + tiny;
  for (int round = 0; round < 64; ++round) {
    const double width = hi - lo;
    if (width <= 2.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    // NT points split [lo, hi] into NT + 1 parts; the first point with the
    // eigenvalue below it bounds the new interval
    const double x = lo + width * (double)(tid + 1) / (double)(NT + 1);
    const unsigned long long mask = __ballot(lz_sturm(d, e2, m, x, pivmin) >= 1);
    if (lane == 0) first_w[wid] = mask ? __ffsll((long long)mask) - 1 + 64 * wid : NT;
    __syncthreads();
    int first = NT;
    for (int w = 0; w < NW; ++w) first = min(first, first_w[w]);
    __syncthreads();
    const double nlo = first == 0 ? lo : lo + width * (double)first / (double)(NT + 1);
    const double nhi = first == NT ? hi : lo + width * (double)(first + 1) / (double)(NT + 1);
    lo = nlo;
    hi = nhi;
  }
  if (tid == 0) {
    const double theta = 0.5 * (lo + hi);
    const bool done = fabs(theta - prev) <= fmax(tol, rtol * fabs(theta)) || m >= max_iter;
    tmp[1] = fabs(theta - prev);
    tmp[0] = theta;
    flags[LZ_LAST] = m;
    if (done) flags[LZ_DONE] = 1;
  }
}

// start of a pass: flags cleared, the stopping rule stored
__global__ void k_lz_init(double* __restrict__ tmp, int* __restrict__ flags, double tol, double rtol, int max_iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    tmp[5] = tol;
    tmp[6] = rtol;
    tmp[7] = (double)max_iter;
    flags[LZ_DONE] = 0;
    flags[LZ_LAST] = 1;
    flags[LZ_BRK] = 0;
  }
}

// the first Ritz value alpha_1, and a first "change" as wide as it
__global__ void k_lz_ritz0(const double* __restrict__ ab, double* __restrict__ tmp) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    tmp[0] = ab[1];
    tmp[1] = fabs(ab[1]) + 1.0;
  }
}

// Pinned host slots for the flags a pass copies back (8 ints each, two
// chunks): carved from slabs that are allocated once and kept until the
// process ends, since pinning and unpinning host memory per plan would
// synchronise the device under every solve.
std::mutex g_lz_mu;
std::vector<int*> g_lz_free;

int* lz_slot_take() {
  std::lock_guard<std::mutex> lk(g_lz_mu);
  if (g_lz_free.empty()) {
    constexpr int SLOTS = 256;
    int* slab = nullptr;
    if (hipHostMalloc((void**)&slab, sizeof(int) * 8 * SLOTS, hipHostMallocDefault) != hipSuccess) return nullptr;
    for (int i = SLOTS - 1; i >= 0; --i) g_lz_free.push_back(slab + 8 * i);
  }
  int* s = g_lz_free.back();
  g_lz_free.pop_back();
  return s;
}

void lz_slot_give(int* s) {
  std::lock_guard<std::mutex> lk(g_lz_mu);
  g_lz_free.push_back(s);
}

}  // namespace

// ---------------------------------------------------------------------------
// plan object
// ---------------------------------------------------------------------------
struct tjsq_plan {
  HostPlan hp;
  int device = 0;
  double* arena = nullptr;       // FP64 operators (sys, env, merged)
  float* arena32 = nullptr;      // FP32 copy (FP32 contractions)
  uint32_t* arenab = nullptr;    // BF16x3 operand words
  uint32_t* xw = nullptr;        // BF16x3 operand words of an apply's x
  double* t64 = nullptr;         // stage-1 scratch
  float* t32 = nullptr;
  Tile* tile1 = nullptr;
  Tile* tile2 = nullptr;
  Seg* seg1 = nullptr;
  Seg* seg2 = nullptr;
  Tile* rtile = nullptr;
  double* part64 = nullptr;      // partial tiles of split stage-2 work
  float* part32 = nullptr;
  // while a Lanczos pass launches its iterations: the pass's "done" flag;
  // every apply kernel of the plan then returns at once after convergence
  const int* gate = nullptr;
  int* lz_host = nullptr;        // pinned copies of the pass flags (two chunks)
  hipEvent_t lz_ev[2] = {nullptr, nullptr};
  // Lanczos workspace
  void* kry = nullptr;           // Krylov basis
  int64_t kry_bytes = 0;
  double* dwork = nullptr;       // partials + dot outputs + coefficients
  float* x32 = nullptr;          // FP32 staging of the mixed solve's applies (x and y)
  float* y32 = nullptr;
  double* x64a = nullptr;
  double* x64b = nullptr;
  // Lanczos iterations captured as hipGraphs, keyed by (precision, basis
  // size); valid while the Krylov basis stays at `graph_kry`
  bool use_graphs = false;
  std::vector<std::pair<int, hipGraphExec_t>> iter_graphs;
  void* graph_kry = nullptr;
  // per-launch HIP events of the two apply stages while profiling is on
  bool profiling = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof;
  // packed BF16x3 apply (tjsq_packed.hpp); used for BF16x3 when `packed`
  bool packed = false;
  tjsq::PackedPlan pk;           // device form of the tables (absolute addresses)
  tjsq::PackedPlan pk_host;      // host form (offsets into the plan's spaces)
  std::vector<void*> pk_owned;   // further buffers the tables address (batch copies)
  // device tables of the set-up kernels (operator packing, hmerge), kept
  // with the plan: freeing them would synchronise the whole device, which
  // stalls plan creation behind other streams' work (a serving mix)
  tjsq::PackJob* pk_ojobs = nullptr;
  PackPiece* pk_opieces = nullptr;
  int pk_nop = 0;
  MergeDst* mg_dsts = nullptr;
  tjsq::MergeTerm* mg_terms = nullptr;
  int mg_ndst = 0;
  uint8_t* pk_ops = nullptr;     // packed operator copies
  uint8_t* pk_x = nullptr;       // packed copies of an apply's x
  uint8_t* pk_t = nullptr;       // stage-1 products
  tjsq::PackJob* pk_xjobs = nullptr;
  PackPiece* pk_xpieces = nullptr;   // the x jobs cut into pieces
  int pk_nxpieces = 0;
  tjsq::PSeg* pk_seg1 = nullptr;
  tjsq::PSeg* pk_seg2 = nullptr;
  tjsq::PTile* pk_tile1 = nullptr;
  tjsq::PTile* pk_tile2 = nullptr;
  tjsq::PReduce* pk_rtile = nullptr;
  float* pk_part = nullptr;
  // the operator-set layouts the plan was compiled for (tjsq_plan_set_ops)
  uint64_t layout_hash = 0;
};

namespace {

// a switch read from the environment (set and not "0")
bool getenv_flag(const char* name) {
  const char* v = std::getenv(name);
  return v && *v && std::strcmp(v, "0") != 0;
}

int grid_for(int64_t n, int threads = 256) {
  int64_t g = (n + threads - 1) / threads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

int upload(tjsq_plan* p, const tjsq_opset& set, int64_t base, hipStream_t st) {
  if (set.ndata == 0) return TJSQ_OK;
  HIPCHK(hipMemcpyAsync(p->arena + base, set.data, sizeof(double) * set.ndata,
                        set.data_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  return TJSQ_OK;
}

// tile shape of the C-ABI -> compiler shape; the kernel has 32x32 and 64x64
// work items and a K chunk of 32
bool to_shape(const tjsq_shape* in, tjsq::Shape& out) {
  out = tjsq::Shape();
  if (!in) return true;
  if ((in->big != 32 && in->big != 64) || in->small != 32 || in->k_split < 32 || in->item_chunks < 1 ||
      in->xcds < 1 || in->xcds > 64 || in->order < 0 || in->order > 1)
    return false;
  out.big = in->big;
  out.small = in->small;
  out.k_split = in->k_split;
  out.item_chunks = in->item_chunks;
  out.xcds = in->xcds;
  out.merge_right = in->merge_right != 0;
  out.order = in->order;
  return true;
}

// device copies of the task tables, FP64 stage-1 scratch and partial tiles
// (plus the FP32 ones when with32)
// FP64 stage-1 scratch and partial tiles, allocated on first FP64 use
int ensure_fp64_scratch(tjsq_plan* p) {
  if (p->t64) return TJSQ_OK;
  const HostPlan& hp = p->hp;
  if (hipMalloc(&p->t64, sizeof(double) * std::max<int64_t>(1, hp.scratch_elems)) != hipSuccess ||
      hipMalloc(&p->part64, sizeof(double) * std::max<int64_t>(1, hp.nslots * 1024)) != hipSuccess) {
    (void)hipFree(p->t64);
    p->t64 = nullptr;
    return fail(TJSQ_ERR_NOMEM, "FP64 scratch allocation failed");
  }
  return TJSQ_OK;
}

int upload_tables(tjsq_plan* p, bool with64, bool with32, hipStream_t st) {
  const HostPlan& hp = p->hp;
  if (with64) {
    int rc = ensure_fp64_scratch(p);
    if (rc) return rc;
  }
  if (with32 && hipMalloc(&p->t32, sizeof(float) * std::max<int64_t>(1, hp.scratch_elems)) != hipSuccess)
    return fail(TJSQ_ERR_NOMEM, "FP32 scratch allocation failed");
  auto put = [&](auto** dptr, const auto& vec) -> int {
    using E = typename std::remove_reference<decltype(vec)>::type::value_type;
    if (hipMalloc((void**)dptr, sizeof(E) * std::max<size_t>(1, vec.size())) != hipSuccess) return TJSQ_ERR_NOMEM;
    if (!vec.empty() && hipMemcpyAsync(*dptr, vec.data(), sizeof(E) * vec.size(), hipMemcpyHostToDevice, st) != hipSuccess)
      return TJSQ_ERR_HIP;
    return TJSQ_OK;
  };
  if (put(&p->tile1, hp.tile1) || put(&p->tile2, hp.tile2) || put(&p->seg1, hp.seg1) || put(&p->seg2, hp.seg2) ||
      put(&p->rtile, hp.rtile))
    return fail(TJSQ_ERR_NOMEM, "task table allocation failed");
  if (with32 && hipMalloc(&p->part32, sizeof(float) * std::max<int64_t>(1, hp.nslots * 1024)) != hipSuccess)
    return fail(TJSQ_ERR_NOMEM, "partial tile allocation failed");
  return TJSQ_OK;
}

// profiling kinds: 0 stage-1 64x64 items, 1 stage-1 32x32, 2 stage-2 64x64,
// 3 stage-2 32x32, 4 the partial-tile reduction
constexpr int NPROF = 5;

template <int MODE, int TSZ, int STAGE>
void launch_items(int n, const Tile* tiles, const Seg* segs, const Spaces<MODE>& S, typename KT<MODE>::T* part,
                  hipStream_t st, const int* gate) {
  hipLaunchKernelGGL((k_tiles<MODE, TSZ, STAGE>), dim3(n), dim3(TileCfg<MODE, TSZ>::NT), 0, st, tiles, segs, S, part,
                     n, gate);
}

template <int MODE, int STAGE>
int launch_stage(const tjsq_plan* p, int n, int n64, const Tile* tiles, const Seg* segs, const Spaces<MODE>& S,
                 typename KT<MODE>::T* part, hipStream_t st) {
  using T = typename KT<MODE>::T;
  auto timed = [&](int kind, auto&& fn) -> int {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (p->profiling) {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      HIPCHK(hipEventRecord(e0, st));
    }
    fn();
    HIPCHK(hipGetLastError());
    if (p->profiling) {
      HIPCHK(hipEventRecord(e1, st));
      const_cast<tjsq_plan*>(p)->prof.push_back({kind, {e0, e1}});
    }
    return TJSQ_OK;
  };
  // 64x64 work items first in the table, then 32x32 ones
  const int k0 = STAGE == 1 ? 0 : 2;
  int rc = TJSQ_OK;
  if (n64) rc = timed(k0, [&] { launch_items<MODE, 64, STAGE>(n64, tiles, segs, S, part, st, p->gate); });
  if (!rc && n > n64)
    rc = timed(k0 + 1, [&] { launch_items<MODE, 32, STAGE>(n - n64, tiles + n64, segs, S, part, st, p->gate); });
  if (!rc && STAGE == 2 && !p->hp.rtile.empty()) {
    const int nr = (int)p->hp.rtile.size();
    rc = timed(4, [&] { hipLaunchKernelGGL(k_reduce<T>, dim3(nr), dim3(256), 0, st, p->rtile, part, S.y, nr, p->gate); });
  }
  return rc;
}

// both stages of an apply over the spaces S
template <int MODE>
int launch_apply(const tjsq_plan* p, const Spaces<MODE>& S, typename KT<MODE>::T* part, hipStream_t st) {
  for (int stage = 1; stage <= 2; ++stage) {
    const int n = stage == 1 ? (int)p->hp.tile1.size() : (int)p->hp.tile2.size();
    if (n == 0) continue;
    const Tile* tiles = stage == 1 ? p->tile1 : p->tile2;
    const Seg* segs = stage == 1 ? p->seg1 : p->seg2;
    const int n64 = stage == 1 ? p->hp.n64_1 : p->hp.n64_2;
    const int rc = stage == 1 ? launch_stage<MODE, 1>(p, n, n64, tiles, segs, S, part, st)
                              : launch_stage<MODE, 2>(p, n, n64, tiles, segs, S, part, st);
    if (rc) return rc;
  }
  return TJSQ_OK;
}

// event-timed launch of one kernel kind while profiling is on
template <typename F>
int timed_launch(const tjsq_plan* p, int kind, hipStream_t st, F&& fn) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (p->profiling) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
  }
  fn();
  HIPCHK(hipGetLastError());
  if (p->profiling) {
    HIPCHK(hipEventRecord(e1, st));
    const_cast<tjsq_plan*>(p)->prof.push_back({kind, {e0, e1}});
  }
  return TJSQ_OK;
}

PSpaces packed_spaces(const float* x, float* y) { return PSpaces{x, y}; }

// packed BF16x3 apply: pack x, stage 1, stage 2, ordered partial sums
// (profiling kinds: 1 the x pack, 0 stage 1, 2 stage 2, 4 the reduction);
// p->pk holds the device form of the tables (counts only are read here)
int packed_apply(const tjsq_plan* p, const float* x, float* y, hipStream_t st) {
  const tjsq::PackedPlan& K = p->pk;
  const PSpaces S = packed_spaces(x, y);
  const int np = p->pk_nxpieces;
  int rc = TJSQ_OK;
  if (np)
    rc = timed_launch(p, 1, st, [&] {
      hipLaunchKernelGGL(k_pack<float>, dim3(std::min(np, 65535)), dim3(PACK_PIECE), 0, st, p->pk_xjobs,
                         p->pk_xpieces, np, x, p->gate);
    });
  const int n1 = (int)K.tile1.size(), n2 = (int)K.tile2.size(), nr = (int)K.rtile.size();
  if (!rc && n1)
    rc = timed_launch(p, 0, st, [&] {
      if (K.quad1)
        hipLaunchKernelGGL(k_packed_group<4>, dim3(n1 / 4), dim3(256), 0, st, p->pk_tile1, p->pk_seg1, n1 / 4,
                           p->gate);
      else
        hipLaunchKernelGGL(k_packed<1>, dim3(n1), dim3(64), 0, st, p->pk_tile1, p->pk_seg1, S, p->pk_part, n1,
                           p->gate);
    });
  if (!rc && n2)
    rc = timed_launch(p, 2, st, [&] {
      hipLaunchKernelGGL(k_packed<2>, dim3(n2), dim3(64), 0, st, p->pk_tile2, p->pk_seg2, S, p->pk_part, n2, p->gate);
    });
  if (!rc && nr)
    rc = timed_launch(p, 4, st, [&] {
      hipLaunchKernelGGL(k_preduce, dim3(nr), dim3(256), 0, st, p->pk_rtile, p->pk_part, y, nr, p->gate);
    });
  return rc;
}

// Device form of one plan's packed tables, appended to D: absolute addresses
// of its operand copies (ops, x) and stage-1 products (t), element offsets
// into x / y shifted by xy, segment indices and partial slots shifted past
// what D already holds.  A plan's own tables use xy = 0; a batch appends
// every member.
void pk_append_device(const tjsq::PackedPlan& K, uint8_t* ops, uint8_t* x, uint8_t* t, int64_t xy,
                      tjsq::PackedPlan& D) {
  const int sb1 = (int)D.seg1.size(), sb2 = (int)D.seg2.size();
  const int64_t slot0 = D.nslots;
  auto fix = [&](tjsq::PSeg g) {
    if (g.l_space() == tjsq::PK_XF) {
      g.l_off += xy;
      return g;
    }
    auto one = [&](int space, int64_t& off, uint32_t& lim) {
      uint8_t* b = space == tjsq::PK_OPS ? ops : space == tjsq::PK_X ? x : t;
      const int64_t n = space == tjsq::PK_OPS ? K.ops_bytes : space == tjsq::PK_X ? K.x_bytes : K.t_bytes;
      lim = (uint32_t)(n - off);
      off = (int64_t)(uintptr_t)(b + off);
    };
    one(g.l_space(), g.l_off, g.l_lim);
    one(g.r_space(), g.r_off, g.r_lim);
    return g;
  };
  for (const tjsq::PSeg& g : K.seg1) D.seg1.push_back(fix(g));
  for (const tjsq::PSeg& g : K.seg2) D.seg2.push_back(fix(g));
  for (tjsq::PTile tl : K.tile1) {
    tl.seg_begin += sb1; tl.seg_end += sb1;
    tl.out_off = (int64_t)(uintptr_t)(t + tl.out_off);
    D.tile1.push_back(tl);
  }
  for (tjsq::PTile tl : K.tile2) {
    tl.seg_begin += sb2; tl.seg_end += sb2;
    tl.out_off += xy;
    if (tl.slot >= 0) tl.slot += (int32_t)slot0;
    D.tile2.push_back(tl);
  }
  for (tjsq::PReduce r : K.rtile) {
    r.out_off += xy;
    r.s0 += (int32_t)slot0; r.s1 += (int32_t)slot0;
    D.rtile.push_back(r);
  }
  for (tjsq::PackJob j : K.x_jobs) {
    j.src_off += xy;
    j.dst_off = (int64_t)(uintptr_t)(x + j.dst_off);
    D.x_jobs.push_back(j);
  }
  D.nslots += K.nslots;
  D.flops_exec += K.flops_exec;
  D.quad1 = K.quad1;   // every plan's tables are compiled with the same shape
}

// device tables of the packed form D (and its partial slots) into p
int pk_upload(tjsq_plan* p, const tjsq::PackedPlan& D, hipStream_t st) {
  auto put = [&](auto** dptr, const auto& vec) -> bool {
    using E = typename std::remove_reference<decltype(vec)>::type::value_type;
    if (hipMalloc((void**)dptr, sizeof(E) * std::max<size_t>(1, vec.size())) != hipSuccess) return false;
    return vec.empty() ||
           hipMemcpyAsync(*dptr, vec.data(), sizeof(E) * vec.size(), hipMemcpyHostToDevice, st) == hipSuccess;
  };
  const std::vector<PackPiece> xp = pack_pieces(D.x_jobs);
  p->pk_nxpieces = (int)xp.size();
  if (hipMalloc(&p->pk_part, sizeof(float) * 4096 * std::max<int64_t>(1, D.nslots)) != hipSuccess ||
      !put(&p->pk_xjobs, D.x_jobs) || !put(&p->pk_xpieces, xp) || !put(&p->pk_seg1, D.seg1) ||
      !put(&p->pk_seg2, D.seg2) ||
      !put(&p->pk_tile1, D.tile1) || !put(&p->pk_tile2, D.tile2) || !put(&p->pk_rtile, D.rtile))
    return fail(TJSQ_ERR_NOMEM, "packed BF16x3 table allocation failed");
  HIPCHK(hipStreamSynchronize(st));   // the host vectors may go after this
  return TJSQ_OK;
}

// the packed operator copies, packed from the FP64 arena (jobs: host form,
// destinations relative to pk_ops)
int pack_ops(tjsq_plan* p, const std::vector<tjsq::PackJob>& jobs, hipStream_t st) {
  if (jobs.empty()) return TJSQ_OK;
  std::vector<tjsq::PackJob> oj = jobs;
  for (tjsq::PackJob& j : oj) j.dst_off = (int64_t)(uintptr_t)(p->pk_ops + j.dst_off);
  if (!p->pk_ojobs) {   // the tables once per plan (the same for every refresh)
    const std::vector<PackPiece> op = pack_pieces(oj);
    if (hipMalloc(&p->pk_ojobs, sizeof(tjsq::PackJob) * oj.size()) != hipSuccess ||
        hipMalloc(&p->pk_opieces, sizeof(PackPiece) * std::max<size_t>(1, op.size())) != hipSuccess) {
      (void)hipFree(p->pk_ojobs);
      p->pk_ojobs = nullptr;
      p->pk_opieces = nullptr;
      return fail(TJSQ_ERR_NOMEM, "packed BF16x3 allocation failed");
    }
    HIPCHK(hipMemcpyAsync(p->pk_ojobs, oj.data(), sizeof(tjsq::PackJob) * oj.size(), hipMemcpyHostToDevice, st));
    if (!op.empty())
      HIPCHK(hipMemcpyAsync(p->pk_opieces, op.data(), sizeof(PackPiece) * op.size(), hipMemcpyHostToDevice, st));
    p->pk_nop = (int)op.size();
  }
  const int np = p->pk_nop;
  if (np)
    hipLaunchKernelGGL(k_pack<double>, dim3(std::min(np, 65535)), dim3(PACK_PIECE), 0, st, p->pk_ojobs, p->pk_opieces,
                       np, p->arena, nullptr);
  const hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(st);
  if (e != hipSuccess || e2 != hipSuccess) return fail(TJSQ_ERR_HIP, "packing the BF16x3 operators failed");
  return TJSQ_OK;
}

// the packed BF16x3 tables and operator copies of a plan; ok = false (and
// the plan keeps the unpacked BF16x3 operands) where the layout does not
// apply or a space would exceed the 2 GiB a buffer resource addresses
int setup_packed(tjsq_plan* p, hipStream_t st, bool& ok) {
  ok = false;
  std::string msg;
  tjsq::PackedPlan K;
  if (tjsq::compile_packed(p->hp, tjsq::PackShape(), K, msg) != TJSQ_OK) return TJSQ_OK;
  const int64_t lim = (int64_t)0x7FFFFFFF - 4096;
  if (K.ops_bytes > lim || K.x_bytes > lim || K.t_bytes > lim) return TJSQ_OK;
  if (hipMalloc(&p->pk_ops, std::max<int64_t>(16, K.ops_bytes)) != hipSuccess ||
      hipMalloc(&p->pk_x, std::max<int64_t>(16, K.x_bytes)) != hipSuccess ||
      hipMalloc(&p->pk_t, std::max<int64_t>(16, K.t_bytes)) != hipSuccess)
    return fail(TJSQ_ERR_NOMEM, "packed BF16x3 allocation failed");
  int rc = pack_ops(p, K.ops_jobs, st);
  if (rc) return rc;
  tjsq::PackedPlan D;
  D.ops_bytes = K.ops_bytes; D.x_bytes = K.x_bytes; D.t_bytes = K.t_bytes;
  pk_append_device(K, p->pk_ops, p->pk_x, p->pk_t, 0, D);
  D.ops_jobs = K.ops_jobs;   // (host form, kept for batches that copy this plan's operators)
  rc = pk_upload(p, D, st);
  if (rc) return rc;
  p->pk_host = std::move(K);
  p->pk = std::move(D);
  ok = true;
  return TJSQ_OK;
}

// hmerge (get_sys.f90:4383-4593) on the device: the merged operators of the
// plan's recipe, summed into the arena in the recipe's accumulation order
int run_merge(tjsq_plan* p, hipStream_t st) {
  const HostPlan& hp = p->hp;
  std::vector<MergeDst> dsts;
  std::vector<tjsq::MergeTerm> terms = hp.merge;
  // group recipe terms by destination (recipes are emitted in accumulation
  // order; stable grouping keeps that order per destination)
  std::stable_sort(terms.begin(), terms.end(),
                   [](const tjsq::MergeTerm& a, const tjsq::MergeTerm& b) { return a.dst_off < b.dst_off; });
  // every merged sub-block (also those without terms) must be zeroed
  const int64_t mstart = hp.merged_base, mend = hp.arena_elems;
  if (mend > mstart) HIPCHK(hipMemsetAsync(p->arena + mstart, 0, sizeof(double) * (mend - mstart), st));
  for (size_t t = 0; t < terms.size();) {
    size_t u = t;
    while (u < terms.size() && terms[u].dst_off == terms[t].dst_off) ++u;
    dsts.push_back({terms[t].dst_off, terms[t].n, (int)t, (int)u});
    t = u;
  }
  if (!dsts.empty()) {
    if (!p->mg_dsts) {   // the tables once per plan (the same for every refresh)
      if (hipMalloc(&p->mg_dsts, sizeof(MergeDst) * dsts.size()) != hipSuccess ||
          hipMalloc(&p->mg_terms, sizeof(tjsq::MergeTerm) * terms.size()) != hipSuccess) {
        (void)hipFree(p->mg_dsts);
        p->mg_dsts = nullptr;
        p->mg_terms = nullptr;
        return fail(TJSQ_ERR_NOMEM, "merge table allocation failed");
      }
      HIPCHK(hipMemcpyAsync(p->mg_dsts, dsts.data(), sizeof(MergeDst) * dsts.size(), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(p->mg_terms, terms.data(), sizeof(tjsq::MergeTerm) * terms.size(), hipMemcpyHostToDevice,
                            st));
      p->mg_ndst = (int)dsts.size();
    }
    int64_t maxn = 1;
    for (auto& d : dsts) maxn = std::max(maxn, d.n);
    dim3 grid((unsigned)std::min<int64_t>((maxn + 255) / 256, 64), (unsigned)std::min<size_t>(dsts.size(), 32768));
    hipLaunchKernelGGL(k_merge, grid, dim3(256), 0, st, p->arena, p->mg_dsts, p->mg_terms, p->mg_ndst);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
  }
  return TJSQ_OK;
}

// the FP32 / BF16x3 operand copies of the FP64 arena a plan holds
int refresh_copies(tjsq_plan* p, hipStream_t st) {
  const HostPlan& hp = p->hp;
  if (p->arena32) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(hp.arena_elems)), dim3(256), 0, st, p->arena, p->arena32,
                       hp.arena_elems);
    HIPCHK(hipGetLastError());
  }
  if (p->packed) {
    int rc = pack_ops(p, p->pk_host.ops_jobs, st);
    if (rc) return rc;
  }
  if (p->arenab) {
    hipLaunchKernelGGL(k_f64_to_split, dim3(grid_for(hp.arena_elems)), dim3(256), 0, st, p->arena, p->arenab,
                       hp.arena_elems);
    HIPCHK(hipGetLastError());
  }
  return TJSQ_OK;
}

// FNV-1a over an operator set's layout (length, header, sub-blocks, offsets)
uint64_t layout_hash(const tjsq_opset& o, uint64_t h) {
  auto mix = [&](const void* d, size_t n) {
    const unsigned char* b = (const unsigned char*)d;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  };
  mix(&o.length, sizeof(int));
  mix(&o.nops, sizeof(int));
  mix(&o.ndata, sizeof(int64_t));
  int64_t nsub = 0;
  for (int k = 0; k < o.nops; ++k) nsub += o.hdr[4 * k + 3];
  mix(o.hdr, sizeof(int) * 4 * (size_t)o.nops);
  mix(o.sub, sizeof(int) * 5 * (size_t)nsub);
  mix(o.off, sizeof(int64_t) * (size_t)nsub);
  return h;
}

// apply of plan p (or a batch's merged plan) in one precision: FP64 on
// double vectors, FP32 / BF16x3 on float vectors
int apply_spaces(const tjsq_plan* p, int precision, const void* x, void* y, hipStream_t st) {
  if (precision == TJSQ_FP64) {
    const double* xd = (const double*)x;
    return launch_apply<MODE_F64>(p, Spaces<MODE_F64>{p->arena, xd, xd, p->t64, (double*)y}, p->part64, st);
  }
  const float* xf = (const float*)x;
  if (precision == TJSQ_FP32)
    return launch_apply<MODE_F32>(p, Spaces<MODE_F32>{p->arena32, xf, xf, p->t32, (float*)y}, p->part32, st);
  if (p->packed) return packed_apply(p, xf, (float*)y, st);
  const int64_t n = p->hp.dim;
  hipLaunchKernelGGL(k_f32_to_split, dim3(grid_for(n)), dim3(256), 0, st, xf, p->xw, n, p->gate);
  HIPCHK(hipGetLastError());
  return launch_apply<MODE_BF16X3>(
      p, Spaces<MODE_BF16X3>{p->arenab, p->xw, xf, (uint32_t*)p->t32, (float*)y}, p->part32, st);
}

// tridiagonal eigenproblem (implicit QL with Wilkinson shifts, Golub & Van
// Loan 8.3); d: diagonal (out: eigenvalues ascending), e: off-diagonal,
// z: n x n eigenvectors (column-major) if wantz
void tridiag_eig(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>* z) {
  if (z) {
    z->assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) (*z)[(size_t)i * n + i] = 1.0;
  }
  e.resize(n + 1, 0.0);
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m != l) {
        if (++iter > 60) break;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (z)
            for (int k = 0; k < n; ++k) {
              double* zc = z->data();
              f = zc[(size_t)(i + 1) * n + k];
              zc[(size_t)(i + 1) * n + k] = s * zc[(size_t)i * n + k] + c * f;
              zc[(size_t)i * n + k] = c * zc[(size_t)i * n + k] - s * f;
            }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  // sort ascending with vectors
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> d2(n);
  std::vector<double> z2;
  if (z) z2.resize((size_t)n * n);
  for (int i = 0; i < n; ++i) {
    d2[i] = d[idx[i]];
    if (z)
      for (int k = 0; k < n; ++k) z2[(size_t)i * n + k] = (*z)[(size_t)idx[i] * n + k];
  }
  d = d2;
  if (z) *z = z2;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* tjsq_last_error(void) { return g_err.c_str(); }

double tjsq_w6js(int j1, int j2, int j3, int l1, int l2, int l3) { return tjsq::w6js(j1, j2, j3, l1, l2, l3); }
double tjsq_w3js(int j1, int j2, int j3, int m1, int m2, int m3) { return tjsq::w3js(j1, j2, j3, m1, m2, m3); }

int tjsq_get_basis(int block_len, int nb, const int* blk, int nsite, const int* site, int* sec, int* sub, int* nsec,
                   int* nsub) {
  (void)block_len;
  if (nb < 0 || nsite < 0 || (!blk && nb) || (!site && nsite) || !sec || !sub || !nsec || !nsub)
    return fail(TJSQ_ERR_ARG, "tjsq_get_basis: null argument");
  std::vector<int> s, q;
  *nsec = tjsq::get_basis(nb, blk, nsite, site, s, q);
  *nsub = (int)(q.size() / 6);
  std::copy(s.begin(), s.end(), sec);
  std::copy(q.begin(), q.end(), sub);
  return TJSQ_OK;
}

int tjsq_get_super_basis(const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int total_up, int total_down, int* wsec,
                         int* nw) {
  if (!sys_bs || !env_bs || !wsec || !nw) return fail(TJSQ_ERR_ARG, "tjsq_get_super_basis: null argument");
  std::vector<int> w;
  *nw = tjsq::get_super_basis(*sys_bs, *env_bs, total_up, total_down, w);
  std::copy(w.begin(), w.end(), wsec);
  return TJSQ_OK;
}

int tjsq_hmerge_tables(const tjsq_lattice* lat, int sys_len, int env_len, int* coup_bs, int* coup_bb, int* inb) {
  if (!lat || !coup_bs || !coup_bb || !inb) return fail(TJSQ_ERR_ARG, "tjsq_hmerge_tables: null argument");
  if (sys_len + env_len + 2 > lat->ny * lat->nx_tables || sys_len < 1 || env_len < 1)
    return fail(TJSQ_ERR_ARG, "tjsq_hmerge_tables: lattice too small");
  tjsq::LatticeTables L;
  tjsq::build_lattice(*lat, L);
  tjsq::lattice_setup(L, lat->ny, sys_len, env_len);
  tjsq::Couplings C;
  tjsq::hmerge_tables(L, sys_len, env_len, C);
  coup_bs[0] = C.coup_bs[1][1];
  coup_bs[1] = C.coup_bs[1][2];
  coup_bs[2] = C.coup_bs[2][1];
  coup_bs[3] = C.coup_bs[2][2];
  coup_bb[0] = C.coup_bb[1];
  coup_bb[1] = C.coup_bb[2];
  for (int i = 0; i <= sys_len + env_len + 2; ++i) inb[i] = C.inb[i];
  return TJSQ_OK;
}

int tjsq_active_sites(const tjsq_lattice* lat, int len) {
  if (!lat || len < 1 || len > lat->ny * lat->nx_tables - 2) return fail(TJSQ_ERR_ARG, "tjsq_active_sites: bad length");
  tjsq::LatticeTables L;
  tjsq::build_lattice(*lat, L);
  return L.nleg11[len];
}

int tjsq_plan_create(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys, const tjsq_opset* env,
                     const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int nw, const int* wsec, int precision_mask,
                     void* stream, tjsq_plan** plan) {
  return tjsq_plan_create_ex(lat, sys_len, env_len, sys, env, sys_bs, env_bs, nw, wsec, precision_mask, nullptr, stream,
                             plan);
}

int tjsq_plan_create_ex(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys,
                        const tjsq_opset* env, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int nw,
                        const int* wsec, int precision_mask, const tjsq_shape* shape, void* stream, tjsq_plan** plan) {
  if (!lat || !sys || !env || !sys_bs || !env_bs || !wsec || !plan || nw < 1)
    return fail(TJSQ_ERR_ARG, "tjsq_plan_create: null argument");
  *plan = nullptr;
  tjsq::Shape sh;
  if (!to_shape(shape, sh)) return fail(TJSQ_ERR_ARG, "tjsq_plan_create: unsupported tile shape");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TJSQ_ERR_NODEV, "no GPU visible");
  tjsq_plan* p = new tjsq_plan();
  std::string msg;
  int rc = tjsq::build_host_plan(*lat, sys_len, env_len, *sys, *env, *sys_bs, *env_bs, nw, wsec, sh, p->hp, msg);
  if (rc != TJSQ_OK) {
    delete p;
    return fail(rc, "tjsq_plan_create: " + msg);
  }
  hipStream_t st = (hipStream_t)stream;
  HostPlan& hp = p->hp;
  auto cleanup = [&](int code, const std::string& m) {
    tjsq_plan_destroy(p);
    return fail(code, m);
  };
  (void)hipGetDevice(&p->device);
  p->layout_hash = layout_hash(*env, layout_hash(*sys, 1469598103934665603ull));
  if (hipMalloc(&p->arena, sizeof(double) * std::max<int64_t>(1, hp.arena_elems)) != hipSuccess)
    return cleanup(TJSQ_ERR_NOMEM, "operator arena allocation failed");
  if ((rc = upload(p, *sys, hp.sys_base, st)) || (rc = upload(p, *env, hp.env_base, st))) {
    tjsq_plan_destroy(p);
    return rc;
  }
  if ((rc = run_merge(p, st))) {
    tjsq_plan_destroy(p);
    return rc;
  }
  if (precision_mask & (1 << TJSQ_FP32)) {
    if (hipMalloc(&p->arena32, sizeof(float) * std::max<int64_t>(1, hp.arena_elems)) != hipSuccess)
      return cleanup(TJSQ_ERR_NOMEM, "FP32 operator arena allocation failed");
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(hp.arena_elems)), dim3(256), 0, st, p->arena, p->arena32,
                       hp.arena_elems);
    HIPCHK(hipGetLastError());
  }
  if (precision_mask & (1 << TJSQ_BF16X3)) {
    bool ok = false;
    if (!getenv_flag("TJSQ_UNPACKED_BF16X3") && (rc = setup_packed(p, st, ok))) {
      tjsq_plan_destroy(p);
      return rc;
    }
    p->packed = ok;
    hp.stats.packed = ok ? 1 : 0;
    hp.stats.flops_packed = ok ? p->pk.flops_exec : 0.0;
    if (!ok) {
      if (hipMalloc(&p->arenab, sizeof(uint32_t) * std::max<int64_t>(1, hp.arena_elems)) != hipSuccess ||
          hipMalloc(&p->xw, sizeof(uint32_t) * std::max<int64_t>(1, hp.dim)) != hipSuccess)
        return cleanup(TJSQ_ERR_NOMEM, "BF16x3 operand allocation failed");
      hipLaunchKernelGGL(k_f64_to_split, dim3(grid_for(hp.arena_elems)), dim3(256), 0, st, p->arena, p->arenab,
                         hp.arena_elems);
      HIPCHK(hipGetLastError());
    }
  }
  if ((rc = upload_tables(p, (precision_mask & 1) != 0, p->arena32 || p->arenab, st))) {
    tjsq_plan_destroy(p);
    return rc;
  }
  HIPCHK(hipStreamSynchronize(st));
  *plan = p;
  return TJSQ_OK;
}

int tjsq_plan_set_ops(tjsq_plan* p, const tjsq_opset* sys, const tjsq_opset* env, void* stream) {
  if (!p || !sys || !env) return fail(TJSQ_ERR_ARG, "tjsq_plan_set_ops: null argument");
  if (layout_hash(*env, layout_hash(*sys, 1469598103934665603ull)) != p->layout_hash)
    return fail(TJSQ_ERR_ARG, "tjsq_plan_set_ops: the operator layouts differ from the plan's");
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = upload(p, *sys, p->hp.sys_base, st)) || (rc = upload(p, *env, p->hp.env_base, st))) return rc;
  if ((rc = run_merge(p, st)) || (rc = refresh_copies(p, st))) return rc;
  HIPCHK(hipStreamSynchronize(st));
  return TJSQ_OK;
}

struct tjsq_hostplan {
  HostPlan hp;
  bool have_pk = false;
  int pk_rc = TJSQ_OK;
  tjsq::PackedPlan pk;
};

int tjsq_hostplan_create(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys,
                         const tjsq_opset* env, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int nw,
                         const int* wsec, tjsq_hostplan** out) {
  return tjsq_hostplan_create_ex(lat, sys_len, env_len, sys, env, sys_bs, env_bs, nw, wsec, nullptr, out);
}

int tjsq_hostplan_create_ex(const tjsq_lattice* lat, int sys_len, int env_len, const tjsq_opset* sys,
                            const tjsq_opset* env, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, int nw,
                            const int* wsec, const tjsq_shape* shape, tjsq_hostplan** out) {
  if (!lat || !sys || !env || !sys_bs || !env_bs || !wsec || !out || nw < 1)
    return fail(TJSQ_ERR_ARG, "tjsq_hostplan_create: null argument");
  *out = nullptr;
  tjsq::Shape sh;
  if (!to_shape(shape, sh)) return fail(TJSQ_ERR_ARG, "tjsq_hostplan_create: unsupported tile shape");
  tjsq_hostplan* h = new tjsq_hostplan();
  std::string msg;
  const int rc = tjsq::build_host_plan(*lat, sys_len, env_len, *sys, *env, *sys_bs, *env_bs, nw, wsec, sh, h->hp, msg);
  if (rc != TJSQ_OK) {
    delete h;
    return fail(rc, "tjsq_hostplan_create: " + msg);
  }
  *out = h;
  return TJSQ_OK;
}

int tjsq_hostplan_sizes(const tjsq_hostplan* h, int64_t* sizes) {
  if (!h || !sizes) return fail(TJSQ_ERR_ARG, "tjsq_hostplan_sizes: null argument");
  const HostPlan& hp = h->hp;
  const int64_t v[] = {(int64_t)hp.seg1.size(), (int64_t)hp.tile1.size(), (int64_t)hp.seg2.size(),
                       (int64_t)hp.tile2.size(), (int64_t)hp.merge.size(), hp.arena_elems, hp.scratch_elems, hp.dim,
                       hp.sys_base, hp.env_base, hp.merged_base, (int64_t)hp.rtile.size(), hp.nslots, hp.n64_1, hp.n64_2};
  std::copy(v, v + 15, sizes);
  return TJSQ_OK;
}

int tjsq_hostplan_tables(const tjsq_hostplan* h, void* seg1, void* tile1, void* seg2, void* tile2, void* merge,
                         void* rtile) {
  if (!h) return fail(TJSQ_ERR_ARG, "tjsq_hostplan_tables: null argument");
  const HostPlan& hp = h->hp;
  if (seg1) std::memcpy(seg1, hp.seg1.data(), sizeof(Seg) * hp.seg1.size());
  if (tile1) std::memcpy(tile1, hp.tile1.data(), sizeof(Tile) * hp.tile1.size());
  if (seg2) std::memcpy(seg2, hp.seg2.data(), sizeof(Seg) * hp.seg2.size());
  if (tile2) std::memcpy(tile2, hp.tile2.data(), sizeof(Tile) * hp.tile2.size());
  if (merge) std::memcpy(merge, hp.merge.data(), sizeof(tjsq::MergeTerm) * hp.merge.size());
  if (rtile) std::memcpy(rtile, hp.rtile.data(), sizeof(Tile) * hp.rtile.size());
  return TJSQ_OK;
}

int tjsq_hostplan_stats(const tjsq_hostplan* h, tjsq_plan_stats* stats) {
  if (!h || !stats) return fail(TJSQ_ERR_ARG, "tjsq_hostplan_stats: null argument");
  *stats = h->hp.stats;
  return TJSQ_OK;
}

int tjsq_hostplan_destroy(tjsq_hostplan* h) {
  delete h;
  return TJSQ_OK;
}

int tjsq_hostplan_packed(const tjsq_hostplan* hc, int64_t* sizes, void* ops_jobs, void* x_jobs, void* seg1,
                         void* tile1, void* seg2, void* tile2, void* rtile) {
  if (!hc || !sizes) return fail(TJSQ_ERR_ARG, "tjsq_hostplan_packed: null argument");
  tjsq_hostplan* h = const_cast<tjsq_hostplan*>(hc);
  if (!h->have_pk) {
    std::string msg;
    h->pk_rc = tjsq::compile_packed(h->hp, tjsq::PackShape(), h->pk, msg);
    h->have_pk = true;
    if (h->pk_rc) return fail(h->pk_rc, "tjsq_hostplan_packed: " + msg);
  }
  if (h->pk_rc) return fail(h->pk_rc, "tjsq_hostplan_packed: layout does not apply");
  const tjsq::PackedPlan& K = h->pk;
  const int64_t v[] = {(int64_t)K.ops_jobs.size(), (int64_t)K.x_jobs.size(), (int64_t)K.seg1.size(),
                       (int64_t)K.tile1.size(), (int64_t)K.seg2.size(), (int64_t)K.tile2.size(),
                       (int64_t)K.rtile.size(), K.nslots, K.ops_bytes, K.x_bytes, K.t_bytes};
  std::copy(v, v + 11, sizes);
  if (ops_jobs) std::memcpy(ops_jobs, K.ops_jobs.data(), sizeof(tjsq::PackJob) * K.ops_jobs.size());
  if (x_jobs) std::memcpy(x_jobs, K.x_jobs.data(), sizeof(tjsq::PackJob) * K.x_jobs.size());
  if (seg1) std::memcpy(seg1, K.seg1.data(), sizeof(tjsq::PSeg) * K.seg1.size());
  if (tile1) std::memcpy(tile1, K.tile1.data(), sizeof(tjsq::PTile) * K.tile1.size());
  if (seg2) std::memcpy(seg2, K.seg2.data(), sizeof(tjsq::PSeg) * K.seg2.size());
  if (tile2) std::memcpy(tile2, K.tile2.data(), sizeof(tjsq::PTile) * K.tile2.size());
  if (rtile) std::memcpy(rtile, K.rtile.data(), sizeof(tjsq::PReduce) * K.rtile.size());
  return TJSQ_OK;
}

int tjsq_plan_reshape(tjsq_plan* p, const tjsq_shape* shape, void* stream) {
  if (!p) return fail(TJSQ_ERR_ARG, "tjsq_plan_reshape: null plan");
  tjsq::Shape sh;
  if (!to_shape(shape, sh)) return fail(TJSQ_ERR_ARG, "tjsq_plan_reshape: unsupported tile shape");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipDeviceSynchronize());   // no launch may still read the old tables
  // the contribution list does not depend on the shape: recompile the tables
  const tjsq_plan_stats keep = p->hp.stats;
  int rc = tjsq::compile_program(p->hp, sh);
  if (rc) return fail(rc, "tjsq_plan_reshape: compilation failed");
  p->hp.stats.dim = keep.dim;
  p->hp.stats.bytes_fp64 = keep.bytes_fp64;
  void* bufs[] = {p->tile1, p->tile2, p->seg1, p->seg2, p->rtile, p->t64, p->t32, p->part64, p->part32};
  for (void* b : bufs) (void)hipFree(b);
  p->tile1 = p->tile2 = nullptr; p->seg1 = p->seg2 = nullptr; p->rtile = nullptr;
  p->t64 = nullptr; p->t32 = nullptr; p->part64 = nullptr; p->part32 = nullptr;
  if ((rc = upload_tables(p, false, p->arena32 || p->arenab, st))) return rc;
  HIPCHK(hipStreamSynchronize(st));
  return TJSQ_OK;
}

int tjsq_plan_stats_get(const tjsq_plan* plan, tjsq_plan_stats* stats) {
  if (!plan || !stats) return fail(TJSQ_ERR_ARG, "tjsq_plan_stats_get: null argument");
  *stats = plan->hp.stats;
  return TJSQ_OK;
}

int tjsq_plan_destroy(tjsq_plan* p) {
  if (!p) return TJSQ_OK;
  for (auto& g : p->iter_graphs) (void)hipGraphExecDestroy(g.second);
  p->iter_graphs.clear();
  for (auto& e : p->lz_ev)
    if (e) (void)hipEventDestroy(e);
  if (p->lz_host) lz_slot_give(p->lz_host);
  (void)hipFree(p->arena);
  (void)hipFree(p->arena32);
  (void)hipFree(p->arenab);
  (void)hipFree(p->xw);
  (void)hipFree(p->t64);
  (void)hipFree(p->t32);
  (void)hipFree(p->tile1);
  (void)hipFree(p->tile2);
  (void)hipFree(p->seg1);
  (void)hipFree(p->seg2);
  (void)hipFree(p->rtile);
  (void)hipFree(p->part64);
  (void)hipFree(p->part32);
  (void)hipFree(p->kry);
  (void)hipFree(p->dwork);
  (void)hipFree(p->x32);
  (void)hipFree(p->y32);
  (void)hipFree(p->x64a);
  (void)hipFree(p->x64b);
  (void)hipFree(p->pk_ops);
  (void)hipFree(p->pk_x);
  (void)hipFree(p->pk_t);
  (void)hipFree(p->pk_xjobs);
  (void)hipFree(p->pk_xpieces);
  (void)hipFree(p->pk_seg1);
  (void)hipFree(p->pk_seg2);
  (void)hipFree(p->pk_tile1);
  (void)hipFree(p->pk_tile2);
  (void)hipFree(p->pk_rtile);
  (void)hipFree(p->pk_part);
  for (void* b : p->pk_owned) (void)hipFree(b);
  (void)hipFree(p->pk_ojobs);
  (void)hipFree(p->pk_opieces);
  (void)hipFree(p->mg_dsts);
  (void)hipFree(p->mg_terms);
  delete p;
  return TJSQ_OK;
}

int tjsq_plan_graphs(tjsq_plan* p, int enable) {
  if (!p) return fail(TJSQ_ERR_ARG, "tjsq_plan_graphs: null plan");
  p->use_graphs = enable != 0;
  return TJSQ_OK;
}

int tjsq_plan_profile(tjsq_plan* p, int enable) {
  if (!p) return fail(TJSQ_ERR_ARG, "tjsq_plan_profile: null plan");
  p->profiling = enable != 0;
  return TJSQ_OK;
}

int tjsq_plan_profile_read(tjsq_plan* p, double* ms, int64_t* launches) {
  if (!p || !ms || !launches) return fail(TJSQ_ERR_ARG, "tjsq_plan_profile_read: null argument");
  for (int k = 0; k < NPROF; ++k) {
    ms[k] = 0.0;
    launches[k] = 0;
  }
  for (auto& r : p->prof) {
    HIPCHK(hipEventSynchronize(r.second.second));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, r.second.first, r.second.second));
    ms[r.first] += t;
    launches[r.first] += 1;
    (void)hipEventDestroy(r.second.first);
    (void)hipEventDestroy(r.second.second);
  }
  p->prof.clear();
  return TJSQ_OK;
}

int tjsq_heff_apply(tjsq_plan* p, const double* x, double* y, void* stream) {
  if (!p || !x || !y) return fail(TJSQ_ERR_ARG, "tjsq_heff_apply: null argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = ensure_fp64_scratch(p);
  if (rc) return rc;
  return apply_spaces(p, TJSQ_FP64, x, y, st);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// H_eff apply that allocates nothing: the stage-1 products and the partial
// tiles of split work items live in one file-scope device array, `scratch`
// (TJSQ_NOALLOC_SCRATCH_BYTES, held by every GPU the library runs on until
// the process ends).  In host code `scratch` names a host-side stand-in, not
// a device address: its device address is looked up once per device with
// hipGetSymbolAddress (that device current) and only that address reaches
// kernels.  Calls on one GPU share the array: calls that overlap on one GPU
// race.  Nothing in it needs clearing: every stage-1 product and every
// partial tile is written in full before it is read.
// ---------------------------------------------------------------------------
namespace {

static __device__ __align__(256) unsigned char scratch[TJSQ_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_addr;   // device address of `scratch` per device

int scratch_address(void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    void* a = nullptr;
    HIPCHK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    if (!a) return fail(TJSQ_ERR_HIP, "tjsq_heff_apply_noalloc: no device address for the scratch array");
    g_scratch_addr[dev] = a;
  }
  *out = g_scratch_addr[dev];
  return TJSQ_OK;
}

}  // namespace

extern "C" {

int tjsq_heff_apply_noalloc(tjsq_plan* p, int precision, const void* x, void* y, void* stream) {
  if (!p || !x || !y) return fail(TJSQ_ERR_ARG, "tjsq_heff_apply_noalloc: null argument");
  if (precision != TJSQ_FP64 && precision != TJSQ_FP32)
    return fail(TJSQ_ERR_ARG, "tjsq_heff_apply_noalloc: precision must be FP64 or FP32 (BF16x3 applies read "
                              "their packed copies through the plan's own buffers)");
  if (precision == TJSQ_FP32 && !p->arena32)
    return fail(TJSQ_ERR_ARG, "tjsq_heff_apply_noalloc: plan was created without FP32 operands");
  if (p->profiling) return fail(TJSQ_ERR_ARG, "tjsq_heff_apply_noalloc: profiling records events; switch it off");
  const HostPlan& hp = p->hp;
  const int64_t w = precision == TJSQ_FP64 ? 8 : 4;
  // stage-1 products, then the partial tiles (units of 32 x 32 elements),
  // each start 256-byte aligned
  const int64_t t_bytes = (w * std::max<int64_t>(1, hp.scratch_elems) + 255) / 256 * 256;
  const int64_t need = t_bytes + w * std::max<int64_t>(1, hp.nslots) * 1024;
  if (need > (int64_t)TJSQ_NOALLOC_SCRATCH_BYTES)
    return fail(TJSQ_ERR_ARG, "tjsq_heff_apply_noalloc: the plan needs " + std::to_string(need) +
                                  " bytes of scratch, the static scratch holds " +
                                  std::to_string((int64_t)TJSQ_NOALLOC_SCRATCH_BYTES));
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  uint8_t* b = (uint8_t*)base;
  if (precision == TJSQ_FP64) {
    const double* xd = (const double*)x;
    return launch_apply<MODE_F64>(p, Spaces<MODE_F64>{p->arena, xd, xd, (double*)b, (double*)y},
                                  (double*)(b + t_bytes), st);
  }
  const float* xf = (const float*)x;
  return launch_apply<MODE_F32>(p, Spaces<MODE_F32>{p->arena32, xf, xf, (float*)b, (float*)y},
                                (float*)(b + t_bytes), st);
}

int tjsq_heff_apply_f32(tjsq_plan* p, int precision, const float* x, float* y, void* stream) {
  if (!p || !x || !y) return fail(TJSQ_ERR_ARG, "tjsq_heff_apply_f32: null argument");
  if (precision != TJSQ_FP32 && precision != TJSQ_BF16X3) return fail(TJSQ_ERR_ARG, "precision must be FP32 or BF16X3");
  if (precision == TJSQ_FP32 && !p->arena32) return fail(TJSQ_ERR_ARG, "plan was created without FP32 operands");
  if (precision == TJSQ_BF16X3 && !p->arenab && !p->packed)
    return fail(TJSQ_ERR_ARG, "plan was created without BF16x3 operands");
  return apply_spaces(p, precision, x, y, (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Lanczos (lanczos_memo_new, eigen.f90:2-175, levels = 1)
// ---------------------------------------------------------------------------
namespace {

template <typename T>
struct Krylov {
  tjsq_plan* p;
  hipStream_t st;
  int64_t n;
  T* V;              // (max_iter + 1) x n
  double* partial;   // 2 * RED_BLOCKS * nv
  double* dots;      // device results (2 * (max_iter + 1))
  double* coef;      // device coefficients
  double* ab;        // device alpha_j at 2j-1, beta_j at 2j (the three-term coefficients are adjacent)
  double* tmp;       // device scratch scalars
  double* W = nullptr;   // FP64 working vector of a solve whose basis is FP32
  std::vector<double> hbuf;

  T* vec(int j) { return V + (int64_t)j * n; }
  // pass state: flags (LZ_*), and V_j . V_j of the stored basis vectors
  int* flags() { return (int*)(tmp + 32); }
  double* vv() { return tmp + 64; }

  // host <- [V_j . w for j < nv] (+ [V_j . V_j] when self)
  int dots_with(int nv, const T* w, bool self, std::vector<double>& out, const T* Vbase = nullptr) {
    const T* base = Vbase ? Vbase : V;
    hipLaunchKernelGGL(k_multidot<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, base, n, nv, w, n, partial,
                       self ? 1 : 0);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_finish_dots, dim3(nv * (self ? 2 : 1)), dim3(RED_THREADS), 0, st, partial, nv, RED_BLOCKS,
                       self ? 1 : 0, dots);
    HIPCHK(hipGetLastError());
    out.resize(self ? 2 * nv : nv);
    HIPCHK(hipMemcpyAsync(out.data(), dots, sizeof(double) * out.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return TJSQ_OK;
  }
  // w = (w - sum_j c_j V_{first+j}) * s
  int combine(T* w, int nv, const std::vector<double>& c, double s, bool into_zero = false, int first = 0) {
    if (nv > 0) HIPCHK(hipMemcpyAsync(coef, c.data(), sizeof(double) * nv, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_combine<T>, dim3(grid_for(n)), dim3(256), 0, st, w, V + (int64_t)first * n, n, nv, coef, n,
                       s, into_zero ? 1 : 0);
    HIPCHK(hipGetLastError());
    return TJSQ_OK;
  }
  int scale(T* w, double s) {
    hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n)), dim3(256), 0, st, w, n, s);
    HIPCHK(hipGetLastError());
    return TJSQ_OK;
  }
  int norm(const T* w, double& out) {
    std::vector<double> d;
    int rc = dots_with(1, w, false, d, w);
    if (rc) return rc;
    out = std::sqrt(d[0]);
    return TJSQ_OK;
  }
  // ---- device-scalar forms (no host round trip) ----
  // out[0] = V_j . w for nv vectors from Vbase (and the self dots after them)
  void dots_dev(int nv, const T* Vbase, const T* w, bool self, double* out) {
    hipLaunchKernelGGL(k_multidot<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, Vbase, n, nv, w, n, partial,
                       self ? 1 : 0);
    hipLaunchKernelGGL(k_finish_dots, dim3(nv * (self ? 2 : 1)), dim3(RED_THREADS), 0, st, partial, nv, RED_BLOCKS,
                       self ? 1 : 0, out);
  }
  // w /= |w|, the norm also stored at nrm
  void normalize_dev(T* w, double* nrm) {
    dots_dev(1, w, w, false, tmp + 8);
    hipLaunchKernelGGL(k_sqrt_to, dim3(1), dim3(64), 0, st, nrm, tmp + 8);
    hipLaunchKernelGGL(k_scale_inv_dev<T>, dim3(grid_for(n)), dim3(256), 0, st, w, n, nrm);
  }
};

int apply_prec(tjsq_plan* p, int precision, const void* x, void* y, hipStream_t st) {
  if (precision == TJSQ_FP64) return tjsq_heff_apply(p, (const double*)x, (double*)y, st);
  return tjsq_heff_apply_f32(p, precision, (const float*)x, (float*)y, st);
}

// y = H x on FP64 vectors (the FP64 solve and the FP64 restarts; a mixed
// solve's basis is FP32 and its apply reads it directly)
int apply_f64(tjsq_plan* p, int precision, const double* x, double* y, hipStream_t st) {
  if (precision != TJSQ_FP64) return fail(TJSQ_ERR_ARG, "an FP64 Krylov basis is applied in FP64 only");
  return tjsq_heff_apply(p, x, y, st);
}

// One Lanczos iteration at basis size `last` (eigen.f90:63-96 with
// Gram_Schmidt_process, block.f90:2515-2551), every scalar on the device, in
// five streaming passes over the vectors:
//   W = H v_{last-1} - alpha_last v_{last-1} - beta_{last-1} v_{last-2} (FP64),
//       beta_last = |W| (a beta at or below tol ends the pass);
//   c_j = (v_j . W / beta) / (v_j . v_j) for j < last;
//   W = W / beta - sum_j c_j v_j, v_last = W / |W| stored in the basis type;
//   v_{last+1} = H v_last, alpha_{last+1} = v_last . v_{last+1};
//   the lowest eigenvalue of the tridiagonal and the stopping rule.
// W is an FP64 working vector for both basis types, so the FP64 and the
// mixed solves run the same kernels.  Once the pass is done every kernel of
// a later iteration, the applies included, returns at once (the flags).
template <typename T>
int iteration_launches(Krylov<T>& K, int precision, int last) {
  const int64_t n = K.n;
  double* W = K.W;
  int* flags = K.flags();
  const dim3 g(RED_BLOCKS), b(RED_THREADS);
  // partial sums: the Gram-Schmidt dots at the start of K.partial, the
  // scalar ones (|W|^2, v . v, alpha) in three RED_BLOCKS regions after them
  double* pw = K.partial + (int64_t)RED_BLOCKS * RED_MAXV;
  double* pv = pw + RED_BLOCKS;
  double* pa = pv + RED_BLOCKS;
  if (last < 2)
    hipLaunchKernelGGL(k_lz_start<T>, g, b, 0, K.st, W, K.vec(last), K.vec(0), n, 1, K.ab + 1, n, pw, flags);
  else
    hipLaunchKernelGGL(k_lz_start<T>, g, b, 0, K.st, W, K.vec(last), K.vec(last - 2), n, 2, K.ab + 2 * last - 2, n,
                       pw, flags);
  hipLaunchKernelGGL(k_lz_dots<T>, dim3(LZ_DOT_BLOCKS, (last + LZ_DOT_GROUP - 1) / LZ_DOT_GROUP), b, 0, K.st, K.V, n,
                     last, W, n, K.partial, flags);
  hipLaunchKernelGGL(k_lz_coefs, dim3(last), b, 0, K.st, K.partial, last, LZ_DOT_BLOCKS, pw, RED_BLOCKS,
                     K.ab + 2 * last, K.vv(), K.coef, K.tmp, flags);
  hipLaunchKernelGGL(k_lz_subtract<T>, g, b, 0, K.st, W, K.V, n, last, K.coef, K.ab + 2 * last, n, pw, flags);
  hipLaunchKernelGGL(k_lz_store<T>, g, b, 0, K.st, W, K.vec(last), n, pw, RED_BLOCKS, pv, flags);
  HIPCHK(hipGetLastError());
  K.p->gate = flags;
  int rc = std::is_same<T, double>::value ? apply_f64(K.p, precision, (const double*)K.vec(last),
                                                      (double*)K.vec(last + 1), K.st)
                                          : apply_prec(K.p, precision, K.vec(last), K.vec(last + 1), K.st);
  K.p->gate = nullptr;
  if (rc) return rc;
  hipLaunchKernelGGL(k_lz_dot2<T>, g, b, 0, K.st, K.vec(last), K.vec(last + 1), n, pa, flags);
  hipLaunchKernelGGL(k_lz_check, dim3(1), dim3(LZ_CHECK_THREADS), 0, K.st, pa, RED_BLOCKS, pv, K.vv() + last, K.ab,
                     last + 1, K.tmp, flags);
  HIPCHK(hipGetLastError());
  return TJSQ_OK;
}

// the iteration at basis size `last`, replayed from a hipGraph captured on
// first use (plans with graphs enabled, on a non-default stream)
template <typename T>
int iteration(Krylov<T>& K, int precision, int last) {
  tjsq_plan* p = K.p;
  if (!p->use_graphs || p->profiling || K.st == nullptr) return iteration_launches(K, precision, last);
  if (p->graph_kry != p->kry) {     // the basis moved: graphs hold its pointers
    for (auto& g : p->iter_graphs) (void)hipGraphExecDestroy(g.second);
    p->iter_graphs.clear();
    p->graph_kry = p->kry;
  }
  const int key = precision * 1024 + last;
  for (auto& g : p->iter_graphs)
    if (g.first == key) {
      HIPCHK(hipGraphLaunch(g.second, K.st));
      return TJSQ_OK;
    }
  HIPCHK(hipStreamBeginCapture(K.st, hipStreamCaptureModeThreadLocal));
  int rc = iteration_launches(K, precision, last);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(K.st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(TJSQ_ERR_HIP, std::string("iteration capture failed: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(TJSQ_ERR_HIP, std::string("graph instantiation failed: ") + hipGetErrorString(e));
  p->iter_graphs.push_back({key, exec});
  HIPCHK(hipGraphLaunch(exec, K.st));
  return TJSQ_OK;
}

// One Lanczos pass from v0 (already in slot 0, normalised).  Returns the
// number of vectors, the Ritz value and the tridiagonal data.  Iterations
// are issued LZ_CHUNK at a time and the host reads the pass's flags once per
// chunk, while the next chunk is already queued; the iterations queued after
// the pass has ended return at once.
constexpr int LZ_CHUNK = 8;

template <typename T>
int lanczos_pass(Krylov<T>& K, int precision, double tol, int max_iter, int& last_out, double& ritz_out,
                 std::vector<double>& alpha, std::vector<double>& beta, int& breakdown, int& applies,
                 double rtol = 0.0, double* theta2 = nullptr) {
  alpha.assign(max_iter + 1, 0.0);
  beta.assign(max_iter + 1, 0.0);
  breakdown = 0;
  tjsq_plan* p = K.p;
  if (precision == TJSQ_FP64) {
    int rc0 = ensure_fp64_scratch(p);   // no allocation may happen inside a capture
    if (rc0) return rc0;
  }
  if (!p->lz_host) {
    if (!(p->lz_host = lz_slot_take())) return fail(TJSQ_ERR_NOMEM, "pinned flag buffer allocation failed");
    for (int i = 0; i < 2; ++i) HIPCHK(hipEventCreateWithFlags(&p->lz_ev[i], hipEventDisableTiming));
  }
  int* flags = K.flags();
  const dim3 g(RED_BLOCKS), b(RED_THREADS);
  hipLaunchKernelGGL(k_lz_init, dim3(1), dim3(64), 0, K.st, K.tmp, flags, tol, rtol, max_iter);
  int rc = std::is_same<T, double>::value
               ? apply_f64(p, precision, (const double*)K.vec(0), (double*)K.vec(1), K.st)
               : apply_prec(p, precision, K.vec(0), K.vec(1), K.st);
  if (rc) return rc;
  ++applies;
  hipLaunchKernelGGL(k_lz_dot2<T>, g, b, 0, K.st, K.vec(0), K.vec(1), K.n, K.partial, flags);
  hipLaunchKernelGGL(k_lz_sum, dim3(1), b, 0, K.st, K.partial, RED_BLOCKS, K.ab + 1, flags);
  hipLaunchKernelGGL(k_lz_dot2<T>, g, b, 0, K.st, K.vec(0), K.vec(0), K.n, K.partial, flags);
  hipLaunchKernelGGL(k_lz_sum, dim3(1), b, 0, K.st, K.partial, RED_BLOCKS, K.vv(), flags);
  hipLaunchKernelGGL(k_lz_ritz0, dim3(1), dim3(64), 0, K.st, K.ab, K.tmp);
  HIPCHK(hipGetLastError());
  int next = 1;        // basis size of the next iteration to issue
  auto issue = [&](int slot) -> int {
    for (int c = 0; c < LZ_CHUNK && next < max_iter; ++c, ++next) {
      int r = iteration(K, precision, next);
      if (r) return r;
    }
    HIPCHK(hipMemcpyAsync(p->lz_host + 4 * slot, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, K.st));
    HIPCHK(hipEventRecord(p->lz_ev[slot], K.st));
    return TJSQ_OK;
  };
  int slot = 0;
  if ((rc = issue(0))) return rc;
  for (;;) {
    const bool more = next < max_iter;
    if (more && (rc = issue(1 - slot))) return rc;
    HIPCHK(hipEventSynchronize(p->lz_ev[slot]));
    if (p->lz_host[4 * slot + LZ_DONE] || !more) break;
    slot = 1 - slot;
  }
  // the pass's tridiagonal, its size and flags (the queued iterations after
  // the end changed nothing)
  std::vector<double> ab(2 * max_iter + 2, 0.0);
  double tmp01[2];
  HIPCHK(hipMemcpyAsync(ab.data(), K.ab, sizeof(double) * ab.size(), hipMemcpyDeviceToHost, K.st));
  HIPCHK(hipMemcpyAsync(tmp01, K.tmp, sizeof(tmp01), hipMemcpyDeviceToHost, K.st));
  HIPCHK(hipMemcpyAsync(p->lz_host, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, K.st));
  HIPCHK(hipStreamSynchronize(K.st));
  const int last = p->lz_host[LZ_LAST];
  breakdown = p->lz_host[LZ_BRK];
  if (last < 1 || last > max_iter) return fail(TJSQ_ERR_HIP, "Lanczos pass state out of range");
  for (int i = 1; i <= last; ++i) alpha[i] = ab[2 * i - 1];
  for (int i = 1; i < last + (breakdown ? 1 : 0) && i <= max_iter; ++i) beta[i] = ab[2 * i];
  applies += last - 1;
  if (theta2) {
    std::vector<double> dia(alpha.begin() + 1, alpha.begin() + last + 1);
    std::vector<double> off(beta.begin() + 1, beta.begin() + last + 1);
    tridiag_eig(last, dia, off, nullptr);
    *theta2 = last > 1 ? dia[1] : 0.0;
  }
  last_out = last;
  ritz_out = tmp01[0];
  return TJSQ_OK;
}

// psi = normalise(sum_i z_i v_i) into out (slot-independent buffer)
template <typename T>
int assemble(Krylov<T>& K, int last, const std::vector<double>& alpha, const std::vector<double>& beta, T* out,
             double& e0) {
  std::vector<double> dia(alpha.begin() + 1, alpha.begin() + last + 1);
  std::vector<double> off(beta.begin() + 1, beta.begin() + last + 1);
  std::vector<double> z;
  tridiag_eig(last, dia, off, &z);
  e0 = dia[0];
  std::vector<double> c(last);
  for (int i = 0; i < last; ++i) c[i] = -z[i];  // combine subtracts
  int rc = K.combine(out, last, c, 1.0, true);
  if (rc) return rc;
  double nrm;
  if ((rc = K.norm(out, nrm))) return rc;
  return K.scale(out, 1.0 / nrm);
}

// psi = normalise(sum_i z_i v_i) in FP64 from an FP32-stored basis
int assemble_f64(Krylov<float>& K, int last, const std::vector<double>& alpha, const std::vector<double>& beta,
                 double* out, double& e0) {
  std::vector<double> dia(alpha.begin() + 1, alpha.begin() + last + 1);
  std::vector<double> off(beta.begin() + 1, beta.begin() + last + 1);
  std::vector<double> z;
  tridiag_eig(last, dia, off, &z);
  e0 = dia[0];
  std::vector<double> c(last);
  for (int i = 0; i < last; ++i) c[i] = -z[i];  // combine subtracts
  HIPCHK(hipMemcpyAsync(K.coef, c.data(), sizeof(double) * last, hipMemcpyHostToDevice, K.st));
  hipLaunchKernelGGL((k_combine<double, float>), dim3(grid_for(K.n)), dim3(256), 0, K.st, out, K.V, K.n, last, K.coef,
                     K.n, 1.0, 1);
  HIPCHK(hipGetLastError());
  Krylov<double> D;
  D.p = K.p; D.st = K.st; D.n = K.n; D.V = out; D.partial = K.partial; D.dots = K.dots; D.coef = K.coef;
  double nrm;
  int rc = D.norm(out, nrm);
  if (rc) return rc;
  return D.scale(out, 1.0 / nrm);
}

template <typename T>
int ensure_krylov(tjsq_plan* p, int max_iter, Krylov<T>& K, hipStream_t st) {
  const int64_t n = p->hp.dim;
  const int64_t need = (int64_t)sizeof(T) * n * (max_iter + 2);
  if (p->kry_bytes < need) {
    (void)hipFree(p->kry);
    p->kry = nullptr;
    p->kry_bytes = 0;
    if (hipMalloc(&p->kry, need) != hipSuccess) return fail(TJSQ_ERR_NOMEM, "Krylov basis allocation failed");
    p->kry_bytes = need;
  }
  if (!p->dwork) {
    if (hipMalloc(&p->dwork, sizeof(double) * (2 * RED_BLOCKS * 256 + 2048)) != hipSuccess)
      return fail(TJSQ_ERR_NOMEM, "workspace allocation failed");
  }
  K.p = p;
  K.st = st;
  K.n = n;
  K.V = (T*)p->kry;
  K.partial = p->dwork;
  K.dots = p->dwork + 2 * RED_BLOCKS * 256;
  K.coef = K.dots + 512;
  K.ab = K.dots + 1024;
  K.tmp = K.ab + 512;
  return TJSQ_OK;
}

}  // namespace

extern "C" int tjsq_lanczos_solve(tjsq_plan* p, double* psi, const tjsq_lanczos_opts* opts,
                                  tjsq_lanczos_result* res, void* stream) {
  if (!p || !psi || !opts || !res) return fail(TJSQ_ERR_ARG, "tjsq_lanczos_solve: null argument");
  static_assert(250 + 1 <= RED_MAXV, "a Gram-Schmidt dot takes up to max_iter + 1 vectors");
  if (opts->max_iter < 2 || opts->max_iter > 250) return fail(TJSQ_ERR_ARG, "max_iter must be in [2, 250]");
  if (opts->precision == TJSQ_FP32 && !p->arena32) return fail(TJSQ_ERR_ARG, "plan has no FP32 operands");
  if (opts->precision == TJSQ_BF16X3 && !p->arenab && !p->packed)
    return fail(TJSQ_ERR_ARG, "plan has no BF16x3 operands");
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = (hipStream_t)stream;
  std::memset(res, 0, sizeof(*res));
  const int64_t n = p->hp.dim;
  int rc;
  std::vector<double> alpha, beta;
  int last = 0, brk = 0, applies = 0;
  double ritz = 0.0, e0 = 0.0, theta2 = 0.0;
  if (!p->x64a) {
    if (hipMalloc(&p->x64a, sizeof(double) * n) != hipSuccess || hipMalloc(&p->x64b, sizeof(double) * n) != hipSuccess)
      return fail(TJSQ_ERR_NOMEM, "residual buffers allocation failed");
  }
  if (opts->precision == TJSQ_FP64) {
    Krylov<double> K;
    if ((rc = ensure_krylov(p, opts->max_iter, K, st))) return rc;
    K.W = p->x64b;
    HIPCHK(hipMemcpyAsync(K.vec(0), psi, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    double nrm;
    if ((rc = K.norm(K.vec(0), nrm))) return rc;
    if ((rc = K.scale(K.vec(0), 1.0 / nrm))) return rc;
    if ((rc = lanczos_pass(K, TJSQ_FP64, opts->tol, opts->max_iter, last, ritz, alpha, beta, brk, applies,
                           opts->rtol > 0 ? opts->rtol : 0.0, &theta2)))
      return rc;
    if ((rc = assemble(K, last, alpha, beta, psi, e0))) return rc;
    res->e0 = e0;
    res->e_ritz = e0;
    res->iterations = last;
  } else {
    // mixed precision: FP32 / BF16x3 contractions on an FP32-stored Krylov
    // basis; each new direction is orthogonalised in FP64 (FP64 dots,
    // coefficients and subtraction into an FP64 working vector) before it is
    // stored, and the Ritz vector is assembled in FP64
    Krylov<float> K;
    if ((rc = ensure_krylov(p, opts->max_iter, K, st))) return rc;
    K.W = p->x64b;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n)), dim3(256), 0, st, psi, K.vec(0), n);
    HIPCHK(hipGetLastError());
    double nrm;
    if ((rc = K.norm(K.vec(0), nrm))) return rc;
    if ((rc = K.scale(K.vec(0), 1.0 / nrm))) return rc;
    const double rtol = opts->rtol > 0 ? opts->rtol : 1e-7;
    if ((rc = lanczos_pass(K, opts->precision, opts->tol, opts->max_iter, last, ritz, alpha, beta, brk, applies, rtol,
                           &theta2)))
      return rc;
    if ((rc = assemble_f64(K, last, alpha, beta, psi, e0))) return rc;
    res->e_ritz = e0;
    res->iterations = last;
  }
  // FP64 Rayleigh quotient and residual |H psi - E psi| (and refinement)
  int refinements = 0;
  for (;;) {
    Krylov<double> K2;
    K2.p = p;
    K2.st = st;
    K2.n = n;
    K2.V = psi;
    K2.partial = p->dwork;
    K2.dots = p->dwork + 2 * RED_BLOCKS * 256;
    K2.coef = K2.dots + 512;
    if ((rc = tjsq_heff_apply(p, psi, p->x64a, st))) return rc;
    ++applies;
    std::vector<double> d;
    if ((rc = K2.dots_with(1, p->x64a, true, d))) return rc;  // <psi,Hpsi>, <psi,psi>
    const double rq = d[0] / d[1];
    std::vector<double> c = {rq};
    if ((rc = K2.combine(p->x64a, 1, c, 1.0))) return rc;    // H psi - rq psi
    double r;
    if ((rc = K2.norm(p->x64a, r))) return rc;
    res->residual = r / std::sqrt(d[1]);
    if (opts->precision != TJSQ_FP64) res->e0 = rq;
    res->gap = theta2 > rq ? theta2 - rq : 0.0;
    res->energy_err_est = res->gap > 0 ? res->residual * res->residual / res->gap : INFINITY;
    const bool good = opts->energy_tol > 0 ? res->energy_err_est <= opts->energy_tol
                                           : res->residual <= opts->refine_tol;
    if (opts->precision == TJSQ_FP64 || good || refinements >= opts->max_refine) break;
    // refinement restart: FP64 Lanczos from the current psi
    Krylov<double> K;
    if ((rc = ensure_krylov(p, opts->max_iter, K, st))) return rc;
    K.W = p->x64b;
    HIPCHK(hipMemcpyAsync(K.vec(0), psi, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    double nrm;
    if ((rc = K.norm(K.vec(0), nrm))) return rc;
    if ((rc = K.scale(K.vec(0), 1.0 / nrm))) return rc;
    int last2 = 0, brk2 = 0;
    if ((rc = lanczos_pass(K, TJSQ_FP64, opts->tol, opts->max_iter, last2, ritz, alpha, beta, brk2, applies, 0.0,
                           &theta2)))
      return rc;
    if ((rc = assemble(K, last2, alpha, beta, psi, e0))) return rc;
    res->iterations += last2;
    res->refine_iterations += last2;
    brk |= brk2;
    ++refinements;
    res->e0 = e0;
    res->e_ritz = e0;
  }
  HIPCHK(hipStreamSynchronize(st));
  res->applies = applies;
  res->refinements = refinements;
  res->breakdown = brk;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return TJSQ_OK;
}

// ---------------------------------------------------------------------------
// host-buffer entry points
// ---------------------------------------------------------------------------
extern "C" int tjsq_heff_apply_host(tjsq_plan* p, const double* x, double* y) {
  if (!p || !x || !y) return fail(TJSQ_ERR_ARG, "tjsq_heff_apply_host: null argument");
  const int64_t n = p->hp.dim;
  if (!p->x64a) {
    if (hipMalloc(&p->x64a, sizeof(double) * n) != hipSuccess || hipMalloc(&p->x64b, sizeof(double) * n) != hipSuccess)
      return fail(TJSQ_ERR_NOMEM, "staging buffer allocation failed");
  }
  HIPCHK(hipMemcpy(p->x64b, x, sizeof(double) * n, hipMemcpyHostToDevice));
  double* yd = nullptr;
  if (hipMalloc(&yd, sizeof(double) * n) != hipSuccess) return fail(TJSQ_ERR_NOMEM, "staging buffer allocation failed");
  int rc = tjsq_heff_apply(p, p->x64b, yd, nullptr);
  if (rc == TJSQ_OK && hipMemcpy(y, yd, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "copy of y to the host failed");
  (void)hipFree(yd);
  return rc;
}

extern "C" int tjsq_lanczos_solve_host(tjsq_plan* p, double* psi, const tjsq_lanczos_opts* opts,
                                       tjsq_lanczos_result* res) {
  if (!p || !psi || !opts || !res) return fail(TJSQ_ERR_ARG, "tjsq_lanczos_solve_host: null argument");
  const int64_t n = p->hp.dim;
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * n) != hipSuccess) return fail(TJSQ_ERR_NOMEM, "psi buffer allocation failed");
  int rc = TJSQ_OK;
  if (hipMemcpy(d, psi, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "copy of psi to the device failed");
  if (rc == TJSQ_OK) rc = tjsq_lanczos_solve(p, d, opts, res, nullptr);
  if (rc == TJSQ_OK && hipMemcpy(psi, d, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "copy of psi to the host failed");
  (void)hipFree(d);
  return rc;
}

// ---------------------------------------------------------------------------
// density matrix, truncation and block-update programs (tjsq_update.hpp)
// ---------------------------------------------------------------------------
struct tjsq_program {
  tjsq_plan* dev = nullptr;      // compiled tables (dev->hp) and their device copies
  tjsq::OutLayout out;
  int64_t x_elems = 0, ops_elems = 0;
  int kind = 0;                  // 0: block update / projection, 1: density matrix, 2: wave transform
  int levels = 1;
  bool uploaded = false;
  double* work = nullptr;        // density: sum of the states
  double* dwork = nullptr;       // density / wave transform: dot partials, norm, coefficients
};

namespace {

int make_program(tjsq::Program& pg, int kind, int levels, tjsq_program** out) {
  tjsq_program* P = new tjsq_program();
  P->dev = new tjsq_plan();
  P->dev->hp = std::move(pg.hp);
  P->out = std::move(pg.out);
  P->x_elems = pg.x_elems;
  P->ops_elems = pg.ops_elems;
  P->kind = kind;
  P->levels = levels;
  *out = P;
  return TJSQ_OK;
}

// y *= 1 / s[0]
__global__ void k_scale_inv(double* __restrict__ y, int64_t n, const double* __restrict__ s) {
  const double f = 1.0 / s[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] *= f;
}

// identity sub-blocks: data[off + r + c n] = (r == c) for an n x n block
__global__ void k_identity(double* __restrict__ a, int64_t n) {
  const int64_t nn = n * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (i % (n + 1) == 0) ? 1.0 : 0.0;
}

// Per-device eigensolver state: NSOLVE streams (default priority), each with
// its own rocBLAS handle, so batched eigensolves of different size classes
// overlap.  Created on first use and kept until the process ends.
constexpr int NSOLVE = 4;
// one sector inside its class buffer: rho (n x n, ld n) at src, padded
// N x N matrix at dst
struct PadJob { int64_t src, dst; int n, N; };
struct SolverState {
  rocblas_handle h[NSOLVE] = {};
  hipStream_t s[NSOLVE] = {};
  hipEvent_t ev[NSOLVE + 1] = {};
  // work buffers of the batched eigensolve, grown on demand and reused: a
  // hipMalloc / hipFree pair per call costs several ms in a process that
  // holds many large allocations (18 ms against 10 ms per call measured)
  double* pad = nullptr;
  double* D = nullptr;
  double* E = nullptr;
  rocblas_int* info = nullptr;
  PadJob* jobs = nullptr;
  int64_t pad_cap = 0, D_cap = 0, E_cap = 0, info_cap = 0, jobs_cap = 0;
};

// grows a device buffer to hold `need` elements (its old contents are not
// kept); false when the allocation fails
template <typename T>
bool grow_buffer(T** p, int64_t* cap, int64_t need) {
  if (need <= *cap) return true;
  (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const int64_t n = std::max<int64_t>(1, need + need / 4);
  if (hipMalloc((void**)p, sizeof(T) * n) != hipSuccess) return false;
  *cap = n;
  return true;
}
std::mutex g_solver_mu;
std::vector<SolverState*> g_solver;

int solver_state(SolverState** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if ((int)g_solver.size() <= dev) g_solver.resize(dev + 1, nullptr);
  if (!g_solver[dev]) {
    SolverState* S = new SolverState();
    for (int i = 0; i < NSOLVE; ++i) {
      HIPCHK(hipStreamCreateWithFlags(&S->s[i], hipStreamNonBlocking));
      if (rocblas_create_handle(&S->h[i]) != rocblas_status_success)
        return fail(TJSQ_ERR_HIP, "rocblas_create_handle failed");
      if (rocblas_set_stream(S->h[i], S->s[i]) != rocblas_status_success)
        return fail(TJSQ_ERR_HIP, "rocblas_set_stream failed");
    }
    for (int i = 0; i <= NSOLVE; ++i) HIPCHK(hipEventCreateWithFlags(&S->ev[i], hipEventDisableTiming));
    g_solver[dev] = S;
  }
  *out = g_solver[dev];
  return TJSQ_OK;
}

// releases one device's solver state; the caller has made that device current
void solver_release(SolverState* S) {
  (void)hipDeviceSynchronize();
  for (int i = 0; i < NSOLVE; ++i) {
    if (S->h[i]) (void)rocblas_destroy_handle(S->h[i]);
    if (S->s[i]) (void)hipStreamDestroy(S->s[i]);
  }
  for (int i = 0; i <= NSOLVE; ++i)
    if (S->ev[i]) (void)hipEventDestroy(S->ev[i]);
  (void)hipFree(S->pad);
  (void)hipFree(S->D);
  (void)hipFree(S->E);
  (void)hipFree(S->info);
  (void)hipFree(S->jobs);
  delete S;
}

// Padded size classes of the batched eigensolver.  A strided-batched
// rocsolver_dsyevd call costs about a n + c n^2 b for b matrices of size n
// (measured on MI355X: a = 0.0225 ms, c = 1.4e-6 ms), a fixed cost per size
// that the size's Householder steps set.  The sectors, sorted by size, are
// cut into at most `groups` contiguous classes (each padded to its largest
// sector, rounded up to 16) minimising the largest class cost; the classes
// run concurrently (the largest on the caller's stream, the others on solver
// streams).  At a synthetic m = 8000 bond (91 sectors up to 360) three
// classes took 13.8 ms against 18.9 ms for classes of multiples of 64; at
// the swept m = 8000 bond (81 sectors, 37 of them 256-797) four classes take
// 30.8 ms against 33.2 ms for three, and five or six, more classes than a
// process has hardware queues, 38.6-41.3 ms.
std::vector<int> eig_classes(const std::vector<int>& dims, int groups) {
  const int n = (int)dims.size();
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return dims[a] < dims[b]; });
  auto pad = [](int d) { return 16 * ((std::max(1, d) + 15) / 16); };
  auto cost = [&](int N, int b) { return 0.0225 * N + 1.4e-6 * (double)N * N * b; };
  // best[g][i]: minimal largest cost of the first i sorted sectors in g classes
  const double INF = 1e300;
  std::vector<std::vector<double>> best(groups + 1, std::vector<double>(n + 1, INF));
  std::vector<std::vector<int>> cut(groups + 1, std::vector<int>(n + 1, 0));
  best[0][0] = 0.0;
  for (int g = 1; g <= groups; ++g)
    for (int i = 1; i <= n; ++i)
      for (int j = 0; j < i; ++j) {
        if (best[g - 1][j] >= INF) continue;
        const double v = std::max(best[g - 1][j], cost(pad(dims[idx[i - 1]]), i - j));
        if (v < best[g][i]) { best[g][i] = v; cut[g][i] = j; }
      }
  int g = 1;
  for (int h = 1; h <= groups; ++h)
    if (best[h][n] < best[g][n]) g = h;
  std::vector<int> cls(n, 0);
  for (int i = n; i > 0 && g > 0; --g) {
    const int j = cut[g][i];
    for (int q = j; q < i; ++q) cls[idx[q]] = pad(dims[idx[i - 1]]);
    i = j;
  }
  return cls;
}


// dst = [[rho, 0], [0, diag(-(2 + 2j))]]: the padding block decouples exactly
// in the Householder tridiagonalisation (its reflectors are identities and the
// coupling column is zero), and its eigenvalues lie below any density-matrix
// spectrum, so the sector's eigenpairs are the last n ascending ones with
// eigenvectors zero on the padding rows.
__global__ void k_pad_in(const double* __restrict__ rho, double* __restrict__ pad, const PadJob* __restrict__ jobs) {
  const PadJob J = jobs[blockIdx.y];
  const int64_t NN = (int64_t)J.N * J.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < NN; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % J.N), c = (int)(e / J.N);
    double v = 0.0;
    if (r < J.n && c < J.n) v = rho[J.src + r + (int64_t)c * J.n];
    else if (r == c) v = -2.0 * (1 + (r - J.n));
    pad[J.dst + e] = v;
  }
}

// the sector's eigenvectors (rows 0..n-1 of the last n columns) back over rho
__global__ void k_pad_out(double* __restrict__ rho, const double* __restrict__ pad, const PadJob* __restrict__ jobs) {
  const PadJob J = jobs[blockIdx.y];
  const int64_t nn = (int64_t)J.n * J.n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % J.n), c = (int)(e / J.n);
    rho[J.src + e] = pad[J.dst + r + (int64_t)(J.N - J.n + c) * J.N];
  }
}

void fill_info(const tjsq::TruncSelection& sel, tjsq_trunc_info* info) {
  if (!info) return;
  info->num_state = sel.num_state;
  info->dim = sel.dim;
  info->trunc_err = sel.trunc_err;
  info->entropy = sel.entropy;
  info->entropy1 = sel.entropy1;
  info->entropy2 = sel.entropy2;
}

}  // namespace

extern "C" {

int tjsq_block_update_create(const tjsq_lattice* lat, int side, int sys_len, int env_len, const tjsq_opset* pre,
                             const tjsq_basis* basis, const tjsq_opset* trun, tjsq_program** prog) {
  if (!lat || !pre || !basis || !trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_block_update_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_update_program(*lat, side, sys_len, env_len, *pre, *basis, *trun, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_block_update_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_trun_project_create(const tjsq_opset* ops, const int* ndia, const tjsq_opset* trun, tjsq_program** prog) {
  if (!ops || !trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_trun_project_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_projection_program(*ops, ndia, *trun, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_trun_project_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_dia_transport_create(const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis* basis,
                              const tjsq_opset* trun, tjsq_program** prog) {
  if (!basis || !trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_dia_transport_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_dia_transport_program(ops, site, *basis, *trun, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_dia_transport_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_dia_transport_pairs_create(const tjsq_opset* ops, int nprod, const tjsq_opset* site, const tjsq_basis* basis,
                                    const tjsq_opset* trun, tjsq_program** prog) {
  if (!basis || !trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_dia_transport_pairs_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_dia_transport_program(ops, site, nprod, *basis, *trun, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_dia_transport_pairs_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_measure_create(int side, const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis* basis, int nw,
                        const int* wsec, tjsq_program** prog) {
  if (!basis || !wsec || !prog) return fail(TJSQ_ERR_ARG, "tjsq_measure_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_measure_program(side, ops, site, *basis, nw, wsec, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_measure_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_ndia_transport_create(int side, int flag, const tjsq_opset* ops, const tjsq_opset* site,
                               const tjsq_basis* basis, const tjsq_opset* trun, tjsq_program** prog) {
  if (!basis || !trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_ndia_transport_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_ndia_transport_program(side, (char)flag, ops, site, *basis, *trun, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_ndia_transport_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_pair_ndia_create(int side, int flag, const tjsq_opset* ops, const tjsq_opset* site, const tjsq_basis* basis,
                          const tjsq_opset* trun, tjsq_program** prog) {
  if (!ops || !site || !basis || !trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_pair_ndia_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_pair_ndia_program(side, (char)flag, *ops, *site, *basis, *trun, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_pair_ndia_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_cor_ndia_create(int flag, const tjsq_opset* sys_op, const tjsq_opset* env_ops, int env_trans,
                         const tjsq_opset* env_site, int env_site_trans, const tjsq_opset* sys_site, int sys_site_trans,
                         const tjsq_basis* sys_basis, const tjsq_basis* env_basis, int nw, const int* wsec,
                         tjsq_program** prog) {
  if (!sys_op || !sys_basis || !env_basis || !wsec || !prog)
    return fail(TJSQ_ERR_ARG, "tjsq_cor_ndia_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_cor_ndia_program(*sys_op, env_ops, env_trans, env_site, env_site_trans, sys_site,
                                        sys_site_trans, *sys_basis, *env_basis, nw, wsec, (char)flag, tjsq::Shape(),
                                        pg, msg);
  if (rc) return fail(rc, "tjsq_cor_ndia_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_product_ndia_create(int side, int flag, int rank, const tjsq_opset* ops, const tjsq_opset* site,
                             const tjsq_basis* basis, const tjsq_opset* trun, tjsq_program** prog) {
  if (!ops || !site || !basis || !trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_product_ndia_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_product_ndia_program(side, (char)flag, rank, *ops, *site, *basis, *trun, tjsq::Shape(), pg,
                                            msg);
  if (rc) return fail(rc, "tjsq_product_ndia_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_site_cor_create(int kind, int flag, const tjsq_opset* sys_site, int sys_site_trans, const tjsq_opset* env_ops,
                         int env_trans, const tjsq_opset* env_site, int env_site_trans, const tjsq_basis* sys_basis,
                         const tjsq_basis* env_basis, int nw, const int* wsec, tjsq_program** prog) {
  if (!sys_basis || !env_basis || !wsec || !prog) return fail(TJSQ_ERR_ARG, "tjsq_site_cor_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_site_cor_program(kind, sys_site, sys_site_trans, env_ops, env_trans, env_site, env_site_trans,
                                        *sys_basis, *env_basis, nw, wsec, (char)flag, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_site_cor_create: " + msg);
  return make_program(pg, 0, 1, prog);
}

int tjsq_density_create(int side, int len, int nw, const int* wsec, int levels, tjsq_program** prog) {
  if (!wsec || !prog || levels < 1) return fail(TJSQ_ERR_ARG, "tjsq_density_create: bad argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_density_program(side, len, nw, wsec, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_density_create: " + msg);
  return make_program(pg, 1, levels, prog);
}

int tjsq_wave_transform_create(int dir, const tjsq_basis* sys_bs, const tjsq_basis* env_bs, const tjsq_opset* sys_trun,
                               const tjsq_opset* env_trun, int nw_old, const int* wsec_old, int nw_new,
                               const int* wsec_new, tjsq_program** prog) {
  if (!sys_bs || !env_bs || !sys_trun || !env_trun || !prog) return fail(TJSQ_ERR_ARG, "tjsq_wave_transform_create: null argument");
  *prog = nullptr;
  tjsq::Program pg;
  std::string msg;
  int rc = tjsq::build_wave_transform_program(dir, *sys_bs, *env_bs, *sys_trun, *env_trun, nw_old, wsec_old, nw_new,
                                              wsec_new, tjsq::Shape(), pg, msg);
  if (rc) return fail(rc, "tjsq_wave_transform_create: " + msg);
  return make_program(pg, 2, 1, prog);
}

int tjsq_program_sizes(const tjsq_program* P, int64_t* sizes) {
  if (!P || !sizes) return fail(TJSQ_ERR_ARG, "tjsq_program_sizes: null argument");
  const HostPlan& hp = P->dev->hp;
  const int64_t v[14] = {P->out.nops, (int64_t)P->out.off.size(), P->out.ndata, P->x_elems, P->ops_elems,
                         (int64_t)hp.seg1.size(), (int64_t)hp.tile1.size(), (int64_t)hp.seg2.size(),
                         (int64_t)hp.tile2.size(), (int64_t)hp.rtile.size(), hp.nslots, hp.scratch_elems,
                         hp.n64_1, hp.n64_2};
  std::memcpy(sizes, v, sizeof(v));
  return TJSQ_OK;
}

int tjsq_program_layout(const tjsq_program* P, int* hdr, int* sub, int64_t* off) {
  if (!P) return fail(TJSQ_ERR_ARG, "tjsq_program_layout: null argument");
  if (hdr) std::memcpy(hdr, P->out.hdr.data(), sizeof(int) * P->out.hdr.size());
  if (sub) std::memcpy(sub, P->out.sub.data(), sizeof(int) * P->out.sub.size());
  if (off) std::memcpy(off, P->out.off.data(), sizeof(int64_t) * P->out.off.size());
  return TJSQ_OK;
}

int tjsq_program_tables(const tjsq_program* P, void* seg1, void* tile1, void* seg2, void* tile2, void* rtile) {
  if (!P) return fail(TJSQ_ERR_ARG, "tjsq_program_tables: null argument");
  const HostPlan& hp = P->dev->hp;
  if (seg1) std::memcpy(seg1, hp.seg1.data(), sizeof(Seg) * hp.seg1.size());
  if (tile1) std::memcpy(tile1, hp.tile1.data(), sizeof(Tile) * hp.tile1.size());
  if (seg2) std::memcpy(seg2, hp.seg2.data(), sizeof(Seg) * hp.seg2.size());
  if (tile2) std::memcpy(tile2, hp.tile2.data(), sizeof(Tile) * hp.tile2.size());
  if (rtile) std::memcpy(rtile, hp.rtile.data(), sizeof(Tile) * hp.rtile.size());
  return TJSQ_OK;
}

int tjsq_program_stats(const tjsq_program* P, tjsq_plan_stats* stats) {
  if (!P || !stats) return fail(TJSQ_ERR_ARG, "tjsq_program_stats: null argument");
  *stats = P->dev->hp.stats;
  return TJSQ_OK;
}

int tjsq_program_run(tjsq_program* P, const double* ops, const double* x, double* y, void* stream) {
  if (!P || !y || (!x && P->x_elems > 0) || (!ops && P->ops_elems > 0 && P->kind == 0))
    return fail(TJSQ_ERR_ARG, "tjsq_program_run: null argument");
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if (!P->uploaded) {
    (void)hipGetDevice(&P->dev->device);
    if ((rc = upload_tables(P->dev, true, false, st))) return rc;
    P->uploaded = true;
  }
  const double* xin = x;
  double* nrm = nullptr;
  if (P->kind == 1) {
    // Get_density_matrix_*_new (den_mat.f90:84-134): sum the states and
    // normalise; the normalisation is applied to rho as 1 / |psi|^2
    const int64_t n = P->x_elems;
    if (!P->dwork) {
      if (hipMalloc(&P->dwork, sizeof(double) * (2 * RED_BLOCKS + 16 + P->levels)) != hipSuccess)
        return fail(TJSQ_ERR_NOMEM, "density work allocation failed");
      std::vector<double> minus(P->levels, -1.0);
      HIPCHK(hipMemcpy(P->dwork + 2 * RED_BLOCKS + 16, minus.data(), sizeof(double) * P->levels,
                       hipMemcpyHostToDevice));
      if (P->levels > 1 && hipMalloc(&P->work, sizeof(double) * std::max<int64_t>(1, n)) != hipSuccess)
        return fail(TJSQ_ERR_NOMEM, "density work allocation failed");
    }
    if (P->levels > 1) {
      hipLaunchKernelGGL(k_combine<double>, dim3(grid_for(n)), dim3(256), 0, st, P->work, x, n, P->levels,
                         P->dwork + 2 * RED_BLOCKS + 16, n, 1.0, 1);
      xin = P->work;
    }
    nrm = P->dwork + 2 * RED_BLOCKS;
    hipLaunchKernelGGL(k_multidot<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, xin, n, 1, xin, n, P->dwork, 0);
    hipLaunchKernelGGL(k_finish_dots, dim3(1), dim3(RED_THREADS), 0, st, P->dwork, 1, RED_BLOCKS, 0, nrm);
    HIPCHK(hipGetLastError());
  }
  const double* opsp = ops ? ops : xin;
  if ((rc = launch_apply<MODE_F64>(P->dev, Spaces<MODE_F64>{opsp, xin, xin, P->dev->t64, y}, P->dev->part64, st)))
    return rc;
  if (nrm) {
    hipLaunchKernelGGL(k_scale_inv, dim3(grid_for(P->out.ndata)), dim3(256), 0, st, y, P->out.ndata, nrm);
    HIPCHK(hipGetLastError());
  }
  if (P->kind == 2) {
    // wave_normalize (block.f90:2418-2432): y /= |y|, the norm an FP64
    // two-pass reduction
    const int64_t n = P->out.ndata;
    if (!P->dwork && hipMalloc(&P->dwork, sizeof(double) * (2 * RED_BLOCKS + 16)) != hipSuccess)
      return fail(TJSQ_ERR_NOMEM, "wave transform work allocation failed");
    double* sq = P->dwork + 2 * RED_BLOCKS;
    hipLaunchKernelGGL(k_multidot<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, y, n, 1, y, n, P->dwork, 0);
    hipLaunchKernelGGL(k_finish_dots, dim3(1), dim3(RED_THREADS), 0, st, P->dwork, 1, RED_BLOCKS, 0, sq);
    hipLaunchKernelGGL(k_sqrt_to, dim3(1), dim3(64), 0, st, sq + 1, sq);
    hipLaunchKernelGGL(k_scale_inv, dim3(grid_for(n)), dim3(256), 0, st, y, n, sq + 1);
    HIPCHK(hipGetLastError());
  }
  return TJSQ_OK;
}

int tjsq_program_profile(tjsq_program* P, int enable) {
  if (!P) return fail(TJSQ_ERR_ARG, "tjsq_program_profile: null argument");
  return tjsq_plan_profile(P->dev, enable);
}

int tjsq_program_profile_read(tjsq_program* P, double* ms, int64_t* launches) {
  if (!P) return fail(TJSQ_ERR_ARG, "tjsq_program_profile_read: null argument");
  return tjsq_plan_profile_read(P->dev, ms, launches);
}

// host-buffer form of tjsq_program_run for callers that keep their data on
// the host (the Fortran drop-in): ops and x are copied to the device, the
// program runs on the null stream and y comes back
int tjsq_program_run_host(tjsq_program* P, const double* ops, const double* x, double* y) {
  if (!P || !y || (!x && P->x_elems > 0) || (!ops && P->ops_elems > 0 && P->kind != 1))
    return fail(TJSQ_ERR_ARG, "tjsq_program_run_host: null argument");
  // density programs read no operators (ops is ignored for them) and take
  // `levels` states of x_elems values each
  const int64_t no = (ops && P->kind != 1) ? P->ops_elems : 0;
  const int64_t nx = P->x_elems * (P->kind == 1 ? std::max(1, P->levels) : 1), ny = P->out.ndata;
  double *dops = nullptr, *dx = nullptr, *dy = nullptr;
  auto release = [&]() {
    (void)hipFree(dops);
    (void)hipFree(dx);
    (void)hipFree(dy);
  };
  if ((no > 0 && hipMalloc(&dops, sizeof(double) * no) != hipSuccess) ||
      (nx > 0 && hipMalloc(&dx, sizeof(double) * nx) != hipSuccess) ||
      hipMalloc(&dy, sizeof(double) * std::max<int64_t>(1, ny)) != hipSuccess) {
    release();
    return fail(TJSQ_ERR_NOMEM, "tjsq_program_run_host: staging allocation failed");
  }
  int rc = TJSQ_OK;
  if (no > 0 && hipMemcpy(dops, ops, sizeof(double) * no, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "tjsq_program_run_host: copy of ops to the device failed");
  if (rc == TJSQ_OK && nx > 0 && hipMemcpy(dx, x, sizeof(double) * nx, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "tjsq_program_run_host: copy of x to the device failed");
  // y starts cleared: measurement programs leave entries no term reaches
  if (rc == TJSQ_OK && hipMemset(dy, 0, sizeof(double) * std::max<int64_t>(1, ny)) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "tjsq_program_run_host: clearing y failed");
  if (rc == TJSQ_OK) rc = tjsq_program_run(P, dops, dx, dy, nullptr);
  if (rc == TJSQ_OK && ny > 0 && hipMemcpy(y, dy, sizeof(double) * ny, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "tjsq_program_run_host: copy of y to the host failed");
  release();
  return rc;
}

int tjsq_program_destroy(tjsq_program* P) {
  if (!P) return TJSQ_OK;
  tjsq_plan_destroy(P->dev);
  (void)hipFree(P->work);
  (void)hipFree(P->dwork);
  delete P;
  return TJSQ_OK;
}

int tjsq_truncation_select(int nsec, const int* den_sub, const double* eigval, int kept_min, int kept_max,
                           double trun_err, int* counts, double* spectrum, int* order, tjsq_trunc_info* info) {
  tjsq::TruncSelection sel;
  std::string msg;
  int rc = tjsq::select_truncation(nsec, den_sub, eigval, kept_min, kept_max, trun_err, sel, msg);
  if (rc) return fail(rc, "tjsq_truncation_select: " + msg);
  if (counts) std::memcpy(counts, sel.counts.data(), sizeof(int) * sel.counts.size());
  if (spectrum) std::memcpy(spectrum, sel.sorted.data(), sizeof(double) * sel.sorted.size());
  if (order) std::memcpy(order, sel.order.data(), sizeof(int) * sel.order.size());
  fill_info(sel, info);
  return TJSQ_OK;
}

// host-buffer form of tjsq_truncation_operator: den (host) is copied to the
// device, diagonalised there and the eigenvectors copied back over it
int tjsq_truncation_operator_host(int nsec, const int* den_sub, const int64_t* den_off, double* den, int kept_min,
                                  int kept_max, double trun_err, int* trun_sub, int64_t* trun_off, int* trun_num,
                                  double* spectrum, tjsq_trunc_info* info) {
  if (nsec < 1 || !den_sub || !den_off || !den)
    return fail(TJSQ_ERR_ARG, "tjsq_truncation_operator_host: bad argument");
  int64_t n = 0;
  for (int k = 0; k < nsec; ++k)
    n = std::max<int64_t>(n, den_off[k] + (int64_t)std::max(0, den_sub[5 * k + 3]) * std::max(0, den_sub[5 * k + 4]));
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * std::max<int64_t>(1, n)) != hipSuccess)
    return fail(TJSQ_ERR_NOMEM, "tjsq_truncation_operator_host: staging allocation failed");
  int rc = TJSQ_OK;
  if (n > 0 && hipMemcpy(d, den, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "tjsq_truncation_operator_host: copy to the device failed");
  if (rc == TJSQ_OK)
    rc = tjsq_truncation_operator(nsec, den_sub, den_off, d, kept_min, kept_max, trun_err, trun_sub, trun_off,
                                  trun_num, spectrum, info, nullptr);
  if (rc == TJSQ_OK && n > 0 && hipMemcpy(den, d, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(TJSQ_ERR_HIP, "tjsq_truncation_operator_host: copy to the host failed");
  (void)hipFree(d);
  return rc;
}

int tjsq_truncation_operator(int nsec, const int* den_sub, const int64_t* den_off, double* den, int kept_min,
                             int kept_max, double trun_err, int* trun_sub, int64_t* trun_off, int* trun_num,
                             double* spectrum, tjsq_trunc_info* info, void* stream) {
  if (nsec < 1 || !den_sub || !den_off || !den || !trun_sub || !trun_off || !trun_num)
    return fail(TJSQ_ERR_ARG, "tjsq_truncation_operator: bad argument");
  hipStream_t st = (hipStream_t)stream;
  int64_t dim = 0;
  for (int k = 0; k < nsec; ++k) {
    if (den_sub[5 * k + 3] != den_sub[5 * k + 4] || den_sub[5 * k + 4] < 0)
      return fail(TJSQ_ERR_ARG, "tjsq_truncation_operator: density-matrix sector is not square");
    dim += den_sub[5 * k + 4];
  }
  // sectors grouped into at most four padded size classes, one
  // strided-batched DSYEV-like call (ascending eigenvalues, eigenvectors in
  // place; den_mat.f90:192) per class, the classes run concurrently
  std::vector<int> dims, sec_of;
  for (int k = 0; k < nsec; ++k)
    if (den_sub[5 * k + 4] > 0) {
      dims.push_back(den_sub[5 * k + 4]);
      sec_of.push_back(k);
    }
  const std::vector<int> ncls = eig_classes(dims, 4);
  std::vector<int> cls(nsec, 0);            // padded size of each sector's class
  std::vector<std::pair<int, int>> order;   // (class, sector)
  for (size_t q = 0; q < dims.size(); ++q) {
    cls[sec_of[q]] = ncls[q];
    order.push_back({ncls[q], sec_of[q]});
  }
  std::stable_sort(order.begin(), order.end());
  std::vector<PadJob> jobs;
  struct Batch { int N, first, count; int64_t pad_off, d_off; };
  std::vector<Batch> batches;
  int64_t pad_elems = 0, d_elems = 0;
  for (size_t q = 0; q < order.size(); ++q) {
    const int N = order[q].first, k = order[q].second;
    if (batches.empty() || batches.back().N != N) batches.push_back({N, (int)jobs.size(), 0, pad_elems, d_elems});
    jobs.push_back({den_off[k], pad_elems, den_sub[5 * k + 4], N});
    batches.back().count++;
    pad_elems += (int64_t)N * N;
    d_elems += N;
  }
  const int njobs = (int)jobs.size();
  std::vector<double> hD(d_elems);
  std::vector<rocblas_int> hinfo(std::max(1, njobs), 0);
  if (njobs > 0) {
    std::lock_guard<std::mutex> lk(g_solver_mu);
    SolverState* S = nullptr;
    int rc = solver_state(&S);
    if (rc) return rc;
    if (!grow_buffer(&S->pad, &S->pad_cap, pad_elems) || !grow_buffer(&S->D, &S->D_cap, d_elems) ||
        !grow_buffer(&S->E, &S->E_cap, d_elems) || !grow_buffer(&S->info, &S->info_cap, (int64_t)njobs) ||
        !grow_buffer(&S->jobs, &S->jobs_cap, (int64_t)njobs))
      return fail(TJSQ_ERR_NOMEM, "tjsq_truncation_operator: work allocation failed");
    double* pad = S->pad;
    double* D = S->D;
    double* E = S->E;
    rocblas_int* dinfo = S->info;
    PadJob* djobs = S->jobs;
    bool ok = hipMemcpyAsync(djobs, jobs.data(), sizeof(PadJob) * njobs, hipMemcpyHostToDevice, st) == hipSuccess;
    ok = ok && hipEventRecord(S->ev[NSOLVE], st) == hipSuccess;
    for (int i = 0; ok && i < NSOLVE; ++i) ok = hipStreamWaitEvent(S->s[i], S->ev[NSOLVE], 0) == hipSuccess;
    // A class is a few hundred to a thousand small rocsolver launches per
    // call, so issuing the classes from one host thread serialises their
    // launch cost; each solver stream gets its own host thread (three
    // classes: 11.7 -> 9.7 ms on MI355X).  The HIP device is per host
    // thread, so every worker selects the caller's.
    int dev = 0;
    ok = ok && hipGetDevice(&dev) == hipSuccess;
    const int nthr = (int)std::min<size_t>(batches.size(), NSOLVE);
    std::vector<char> tok(std::max(1, nthr), 1);
    // Streams only overlap on different hardware queues, and a process has
    // few (GPU_MAX_HW_QUEUES, 4 by default): streams made after the caller's
    // own ones share them, and two classes on one queue run one after the
    // other (17.7 ms per call measured in a process holding a side stream,
    // 9.8 ms without).  The largest class therefore runs on the caller's
    // stream, which has a queue already, and the others on the first
    // solver streams, the ones made earliest.
    const size_t nb = batches.size();
    auto worker = [&](int t) {
      bool k = hipSetDevice(dev) == hipSuccess;
      hipStream_t sb = t == 0 ? st : S->s[t - 1];
      k = k && rocblas_set_stream(S->h[t], sb) == rocblas_status_success;
      for (size_t i = t; k && i < nb; i += nthr) {
        const Batch& B = batches[nb - 1 - i];   // largest class first
        dim3 grid((unsigned)std::min<int64_t>(((int64_t)B.N * B.N + 255) / 256, 64), (unsigned)B.count);
        hipLaunchKernelGGL(k_pad_in, grid, dim3(256), 0, sb, den, pad, djobs + B.first);
        k = hipGetLastError() == hipSuccess;
        k = k && rocsolver_dsyevd_strided_batched(S->h[t], rocblas_evect_original, rocblas_fill_upper, B.N,
                                                  pad + B.pad_off, B.N, (rocblas_stride)B.N * B.N, D + B.d_off,
                                                  B.N, E + B.d_off, B.N, dinfo + B.first,
                                                  B.count) == rocblas_status_success;
        if (k) {
          hipLaunchKernelGGL(k_pad_out, dim3(grid.x, B.count), dim3(256), 0, sb, den, pad, djobs + B.first);
          k = hipGetLastError() == hipSuccess;
        }
      }
      tok[t] = k;
    };
    if (ok && nthr > 0) {
      std::vector<std::thread> th;
      int started = 1;
      try {
        for (; started < nthr; ++started) th.emplace_back(worker, started);
      } catch (const std::system_error&) {
        // no thread to spare: the remaining streams are issued from this one
      }
      worker(0);
      for (auto& x : th) x.join();
      for (int t = started; t < nthr; ++t) worker(t);
      for (int t = 0; t < nthr; ++t) ok = ok && tok[t];
    }
    for (int i = 0; ok && i < NSOLVE; ++i)
      ok = hipEventRecord(S->ev[i], S->s[i]) == hipSuccess && hipStreamWaitEvent(st, S->ev[i], 0) == hipSuccess;
    ok = ok && hipMemcpyAsync(hD.data(), D, sizeof(double) * d_elems, hipMemcpyDeviceToHost, st) == hipSuccess;
    ok = ok && hipMemcpyAsync(hinfo.data(), dinfo, sizeof(rocblas_int) * njobs, hipMemcpyDeviceToHost, st) == hipSuccess;
    ok = ok && hipStreamSynchronize(st) == hipSuccess;
    if (!ok) return fail(TJSQ_ERR_HIP, "tjsq_truncation_operator: batched eigensolve failed");
  }
  for (int j = 0; j < njobs; ++j)
    if (hinfo[j] != 0) return fail(TJSQ_ERR_HIP, "tjsq_truncation_operator: eigensolver did not converge");
  // eigenvalues per sector in density-matrix order; the padding eigenvalues
  // must come out as placed and below the sector's spectrum
  std::vector<int64_t> dpos(nsec, -1);
  for (size_t b = 0, j = 0; b < batches.size(); ++b)
    for (int q = 0; q < batches[b].count; ++q, ++j)
      dpos[order[j].second] = batches[b].d_off + (int64_t)q * batches[b].N;
  std::vector<double> hev(dim);
  for (int k = 0, p = 0; k < nsec; ++k) {
    const int n = den_sub[5 * k + 4], N = cls[k];
    if (n == 0) continue;
    const double* d = hD.data() + dpos[k];
    for (int j = 0; j < N - n; ++j)
      if (std::fabs(d[j] + 2.0 * (N - n - j)) > 1e-9)
        return fail(TJSQ_ERR_ARG, "tjsq_truncation_operator: density-matrix sector is not positive semi-definite");
    for (int i = 0; i < n; ++i) hev[p++] = d[N - n + i];
  }
  tjsq::TruncSelection sel;
  std::string msg;
  int rc = tjsq::select_truncation(nsec, den_sub, hev.data(), kept_min, kept_max, trun_err, sel, msg);
  if (rc) return fail(rc, "tjsq_truncation_operator: " + msg);
  // den_mat.f90:324-367: sectors with kept states, in density-matrix order;
  // the kept columns are the last `count` eigenvectors, contiguous in place
  int nt = 0;
  for (int k = 0; k < nsec; ++k) {
    const int c = sel.counts[k];
    if (c <= 0) continue;
    const int n = den_sub[5 * k + 4];
    int* t = trun_sub + 5 * nt;
    t[0] = den_sub[5 * k]; t[1] = den_sub[5 * k + 1]; t[2] = 0; t[3] = n; t[4] = c;
    trun_off[nt] = den_off[k] + (int64_t)(n - c) * n;
    ++nt;
  }
  *trun_num = nt;
  if (spectrum) std::memcpy(spectrum, sel.sorted.data(), sizeof(double) * sel.sorted.size());
  fill_info(sel, info);
  return TJSQ_OK;
}

int tjsq_shutdown(void) {
  std::lock_guard<std::mutex> lk(g_solver_mu);
  int cur = 0;
  const bool have = hipGetDevice(&cur) == hipSuccess;
  for (size_t d = 0; d < g_solver.size(); ++d) {
    if (!g_solver[d]) continue;
    if (hipSetDevice((int)d) == hipSuccess) solver_release(g_solver[d]);
    g_solver[d] = nullptr;
  }
  g_solver.clear();
  if (have) (void)hipSetDevice(cur);
  return TJSQ_OK;
}

int tjsq_identity_truncation(const tjsq_basis* basis, int* trun_sub, int64_t* trun_off, int64_t* ndata, double* data,
                             void* stream) {
  if (!basis || !trun_sub || !trun_off || !ndata) return fail(TJSQ_ERR_ARG, "tjsq_identity_truncation: null argument");
  hipStream_t st = (hipStream_t)stream;
  int64_t off = 0;
  for (int k = 0; k < basis->nsec; ++k) {
    const int n = basis->sec[4 * k + 2];
    int* t = trun_sub + 5 * k;
    t[0] = basis->sec[4 * k]; t[1] = basis->sec[4 * k + 1]; t[2] = 0; t[3] = n; t[4] = n;
    trun_off[k] = off;
    if (data && n > 0) {
      hipLaunchKernelGGL(k_identity, dim3(grid_for((int64_t)n * n)), dim3(256), 0, st, data + off, (int64_t)n);
      HIPCHK(hipGetLastError());
    }
    off += (int64_t)n * n;
  }
  *ndata = off;
  return TJSQ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches: many applies of uneven sizes in one grouped launch per stage
// ---------------------------------------------------------------------------
// The members' task tables are merged with their offsets rebased into
// concatenated spaces (operators, x / y, stage-1 scratch, partial tiles), so
// the tile kernel runs the whole batch unchanged.  A member compiled with XCD
// binning (shape xcds > 1) has its work-item lists padded to a multiple of
// the bin count, so its items keep their XCD (item g runs on XCD g % 8) after
// concatenation.
struct tjsq_batch {
  tjsq_plan* dev = nullptr;                  // merged tables, concatenated arenas
  std::vector<int64_t> x_off;                // member offsets in x / y (n + 1 entries)
  hipStream_t cap = nullptr;                 // capture stream (default priority)
  struct Graph { hipGraphExec_t exec = nullptr; hipGraph_t graph = nullptr; const void* x = nullptr; void* y = nullptr; };
  Graph g[3];                                // per precision
  bool use_graph = false;
};

namespace {

void graph_free(tjsq_batch::Graph& G) {
  if (G.exec) (void)hipGraphExecDestroy(G.exec);
  if (G.graph) (void)hipGraphDestroy(G.graph);
  G = tjsq_batch::Graph();
}

template <typename T>
int batch_launch(tjsq_batch* B, int precision, const T* x, T* y, hipStream_t st) {
  return apply_spaces(B->dev, precision, x, y, st);
}

// runs the batch, through a captured hipGraph when enabled (recaptured when
// x / y change: the graph holds the pointers)
template <typename T>
int batch_run(tjsq_batch* B, int precision, const T* x, T* y, hipStream_t st) {
  if (!B->use_graph || B->dev->profiling) return batch_launch<T>(B, precision, x, y, st);
  tjsq_batch::Graph& G = B->g[precision];
  if (!G.exec || G.x != (const void*)x || G.y != (void*)y) {
    graph_free(G);
    if (!B->cap) HIPCHK(hipStreamCreate(&B->cap));
    HIPCHK(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
    int rc = batch_launch<T>(B, precision, x, y, B->cap);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(B->cap, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail(TJSQ_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
    G.graph = graph;
    HIPCHK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    G.x = x;
    G.y = y;
  }
  HIPCHK(hipGraphLaunch(G.exec, st));
  return TJSQ_OK;
}

}  // namespace

extern "C" {

int tjsq_batch_create(tjsq_plan* const* plans, int n, int precision_mask, void* stream, tjsq_batch** out) {
  if (!plans || n < 1 || !out || (precision_mask & 7) == 0) return fail(TJSQ_ERR_ARG, "tjsq_batch_create: bad argument");
  *out = nullptr;
  const bool with64 = (precision_mask & 1) != 0;
  const bool withf = (precision_mask & (1 << TJSQ_FP32)) != 0;
  const bool withb = (precision_mask & (1 << TJSQ_BF16X3)) != 0;
  for (int i = 0; i < n; ++i) {
    if (!plans[i]) return fail(TJSQ_ERR_ARG, "tjsq_batch_create: null plan");
    if (withf && !plans[i]->arena32)
      return fail(TJSQ_ERR_ARG, "tjsq_batch_create: a member plan was created without FP32 operands");
    if (withb && !plans[i]->arenab && !plans[i]->packed)
      return fail(TJSQ_ERR_ARG, "tjsq_batch_create: a member plan was created without BF16x3 operands");
  }
  // BF16x3 on the packed layout when every member has it
  bool all_packed = withb;
  for (int i = 0; i < n; ++i) all_packed = all_packed && plans[i]->packed;
  hipStream_t st = (hipStream_t)stream;
  tjsq_batch* B = new tjsq_batch();
  B->dev = new tjsq_plan();
  (void)hipGetDevice(&B->dev->device);
  HostPlan& M = B->dev->hp;
  std::vector<int64_t> abase(n), tbase(n), sbase(n);
  std::vector<int> s1base(n), s2base(n);
  B->x_off.assign(n + 1, 0);
  int64_t arena = 0, scratch = 0, slots = 0;
  int seg1 = 0, seg2 = 0;
  for (int i = 0; i < n; ++i) {
    const HostPlan& h = plans[i]->hp;
    abase[i] = arena; tbase[i] = scratch; sbase[i] = slots; s1base[i] = seg1; s2base[i] = seg2;
    B->x_off[i + 1] = B->x_off[i] + h.dim;
    arena += h.arena_elems; scratch += h.scratch_elems; slots += h.nslots;
    seg1 += (int)h.seg1.size(); seg2 += (int)h.seg2.size();
  }
  auto rebase_seg = [&](Seg s, int i) {
    auto base = [&](uint8_t sp) -> int64_t {
      return sp == tjsq::SP_OPS ? abase[i] : sp == tjsq::SP_X ? B->x_off[i] : sp == tjsq::SP_T ? tbase[i] : 0;
    };
    s.l_off += base(s.l_space);
    s.r_off += base(s.r_space);
    return s;
  };
  auto rebase_tile = [&](Tile t, int i, int segb) {
    if (t.mt == 0 && t.nt == 0 && t.seg_begin == t.seg_end) return t;   // XCD padding item
    t.out_off += t.out_space == tjsq::SP_T ? tbase[i] : B->x_off[i];
    t.seg_begin += segb;
    t.seg_end += segb;
    if (t.slot >= 0) t.slot += (int32_t)sbase[i];
    return t;
  };
  for (int i = 0; i < n; ++i) {
    const HostPlan& h = plans[i]->hp;
    for (const Seg& s : h.seg1) M.seg1.push_back(rebase_seg(s, i));
    for (const Seg& s : h.seg2) M.seg2.push_back(rebase_seg(s, i));
  }
  for (int part = 0; part < 2; ++part)      // 64x64 items of every member first, then 32x32
    for (int i = 0; i < n; ++i) {
      const HostPlan& h = plans[i]->hp;
      const size_t lo1 = part ? h.n64_1 : 0, hi1 = part ? h.tile1.size() : h.n64_1;
      const size_t lo2 = part ? h.n64_2 : 0, hi2 = part ? h.tile2.size() : h.n64_2;
      for (size_t q = lo1; q < hi1; ++q) M.tile1.push_back(rebase_tile(h.tile1[q], i, s1base[i]));
      for (size_t q = lo2; q < hi2; ++q) M.tile2.push_back(rebase_tile(h.tile2[q], i, s2base[i]));
      if (!part) { M.n64_1 += (int)(hi1 - lo1); M.n64_2 += (int)(hi2 - lo2); }
    }
  for (int i = 0; i < n; ++i)
    for (Tile r : plans[i]->hp.rtile) {
      r.out_off += B->x_off[i];
      r.seg_begin += (int32_t)sbase[i];
      r.seg_end += (int32_t)sbase[i];
      M.rtile.push_back(r);
    }
  M.arena_elems = arena;
  M.scratch_elems = scratch;
  M.nslots = slots;
  M.dim = B->x_off[n];
  for (int i = 0; i < n; ++i) {
    const tjsq_plan_stats& s = plans[i]->hp.stats;
    M.stats.dim += s.dim; M.stats.n_contrib += s.n_contrib; M.stats.n_chain += s.n_chain;
    M.stats.n_stage1 += s.n_stage1; M.stats.n_out_blocks += s.n_out_blocks;
    M.stats.flops_ref += s.flops_ref; M.stats.flops_exec += s.flops_exec; M.stats.bytes_fp64 += s.bytes_fp64;
    M.stats.op_elems += s.op_elems; M.stats.scratch_elems += s.scratch_elems; M.stats.n_slots += s.n_slots;
  }
  M.stats.n_tiles1 = (int64_t)M.tile1.size();
  M.stats.n_tiles2 = (int64_t)M.tile2.size();
  auto cleanup = [&](int code, const std::string& m) {
    tjsq_batch_destroy(B);
    return fail(code, m);
  };
  if (with64 && hipMalloc(&B->dev->arena, sizeof(double) * std::max<int64_t>(1, arena)) != hipSuccess)
    return cleanup(TJSQ_ERR_NOMEM, "batch arena allocation failed");
  if (withf && hipMalloc(&B->dev->arena32, sizeof(float) * std::max<int64_t>(1, arena)) != hipSuccess)
    return cleanup(TJSQ_ERR_NOMEM, "batch FP32 arena allocation failed");
  if (withb && !all_packed &&
      (hipMalloc(&B->dev->arenab, sizeof(uint32_t) * std::max<int64_t>(1, arena)) != hipSuccess ||
       hipMalloc(&B->dev->xw, sizeof(uint32_t) * std::max<int64_t>(1, M.dim)) != hipSuccess))
    return cleanup(TJSQ_ERR_NOMEM, "batch BF16x3 operand allocation failed");
  if (all_packed) {
    // each member's operator copies, x copies and stage-1 products in
    // buffers of the batch's own (members may be destroyed afterwards)
    tjsq::PackedPlan D;
    for (int i = 0; i < n; ++i) {
      const tjsq::PackedPlan& K = plans[i]->pk_host;
      uint8_t *o = nullptr, *x = nullptr, *t = nullptr;
      const bool okm = hipMalloc(&o, std::max<int64_t>(16, K.ops_bytes)) == hipSuccess &&
                       (B->dev->pk_owned.push_back(o), hipMalloc(&x, std::max<int64_t>(16, K.x_bytes)) == hipSuccess) &&
                       (B->dev->pk_owned.push_back(x), hipMalloc(&t, std::max<int64_t>(16, K.t_bytes)) == hipSuccess);
      if (t) B->dev->pk_owned.push_back(t);
      if (!okm) return cleanup(TJSQ_ERR_NOMEM, "batch packed operand allocation failed");
      if (K.ops_bytes && hipMemcpyAsync(o, plans[i]->pk_ops, K.ops_bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return cleanup(TJSQ_ERR_HIP, "batch packed operand copy failed");
      pk_append_device(K, o, x, t, B->x_off[i], D);
    }
    int rc = pk_upload(B->dev, D, st);
    if (rc) {
      tjsq_batch_destroy(B);
      return rc;
    }
    B->dev->pk = std::move(D);
    B->dev->packed = true;
  }
  for (int i = 0; i < n; ++i) {
    const int64_t e = plans[i]->hp.arena_elems;
    if (e == 0) continue;
    if ((with64 && hipMemcpyAsync(B->dev->arena + abase[i], plans[i]->arena, sizeof(double) * e,
                                  hipMemcpyDeviceToDevice, st) != hipSuccess) ||
        (withf && hipMemcpyAsync(B->dev->arena32 + abase[i], plans[i]->arena32, sizeof(float) * e,
                                 hipMemcpyDeviceToDevice, st) != hipSuccess) ||
        (withb && !all_packed && plans[i]->arenab &&
         hipMemcpyAsync(B->dev->arenab + abase[i], plans[i]->arenab, sizeof(uint32_t) * e,
                        hipMemcpyDeviceToDevice, st) != hipSuccess))
      return cleanup(TJSQ_ERR_HIP, "batch arena copy failed");
    if (withb && !all_packed && !plans[i]->arenab) {
      // a packed member keeps no operand words: split its FP64 operators
      hipLaunchKernelGGL(k_f64_to_split, dim3(grid_for(e)), dim3(256), 0, st, plans[i]->arena,
                         B->dev->arenab + abase[i], e);
      if (hipGetLastError() != hipSuccess) return cleanup(TJSQ_ERR_HIP, "batch operand split failed");
    }
  }
  int rc = upload_tables(B->dev, with64, withf || (withb && !all_packed), st);
  if (rc) {
    tjsq_batch_destroy(B);
    return rc;
  }
  HIPCHK(hipStreamSynchronize(st));
  *out = B;
  return TJSQ_OK;
}

int tjsq_batch_offsets(const tjsq_batch* B, int64_t* off) {
  if (!B || !off) return fail(TJSQ_ERR_ARG, "tjsq_batch_offsets: null argument");
  std::memcpy(off, B->x_off.data(), sizeof(int64_t) * B->x_off.size());
  return TJSQ_OK;
}

int tjsq_batch_stats(const tjsq_batch* B, tjsq_plan_stats* stats) {
  if (!B || !stats) return fail(TJSQ_ERR_ARG, "tjsq_batch_stats: null argument");
  *stats = B->dev->hp.stats;
  return TJSQ_OK;
}

int tjsq_batch_graph(tjsq_batch* B, int enable) {
  if (!B) return fail(TJSQ_ERR_ARG, "tjsq_batch_graph: null batch");
  B->use_graph = enable != 0;
  return TJSQ_OK;
}

int tjsq_batch_apply(tjsq_batch* B, const double* x, double* y, void* stream) {
  if (!B || !x || !y) return fail(TJSQ_ERR_ARG, "tjsq_batch_apply: null argument");
  if (!B->dev->arena) return fail(TJSQ_ERR_ARG, "tjsq_batch_apply: the batch was created without FP64");
  return batch_run<double>(B, TJSQ_FP64, x, y, (hipStream_t)stream);
}

int tjsq_batch_apply_f32(tjsq_batch* B, int precision, const float* x, float* y, void* stream) {
  if (!B || !x || !y || (precision != TJSQ_FP32 && precision != TJSQ_BF16X3))
    return fail(TJSQ_ERR_ARG, "tjsq_batch_apply_f32: bad argument");
  if (precision == TJSQ_FP32 && !B->dev->arena32)
    return fail(TJSQ_ERR_ARG, "tjsq_batch_apply_f32: the batch was created without FP32");
  if (precision == TJSQ_BF16X3 && !B->dev->arenab && !B->dev->packed)
    return fail(TJSQ_ERR_ARG, "tjsq_batch_apply_f32: the batch was created without BF16x3");
  return batch_run<float>(B, precision, x, y, (hipStream_t)stream);
}

int tjsq_batch_profile(tjsq_batch* B, int enable) {
  if (!B) return fail(TJSQ_ERR_ARG, "tjsq_batch_profile: null batch");
  return tjsq_plan_profile(B->dev, enable);
}

int tjsq_batch_profile_read(tjsq_batch* B, double* ms, int64_t* launches) {
  if (!B) return fail(TJSQ_ERR_ARG, "tjsq_batch_profile_read: null batch");
  return tjsq_plan_profile_read(B->dev, ms, launches);
}

int tjsq_batch_destroy(tjsq_batch* B) {
  if (!B) return TJSQ_OK;
  for (auto& G : B->g) graph_free(G);
  if (B->cap) (void)hipStreamDestroy(B->cap);
  tjsq_plan_destroy(B->dev);
  delete B;
  return TJSQ_OK;
}

}  // extern "C"
