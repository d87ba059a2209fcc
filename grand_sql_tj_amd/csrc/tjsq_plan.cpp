// tjsq_plan.cpp -- host tables and the contribution compiler for H_eff.
//
// Every routine cites the reference routine it restates
// (DongNingSheng/Grand_SQL_tJ, code/tjsq/<file>.f90:<lines>).  The walk over
// sectors, sub-blocks and spin channels follows the reference's loop and
// search order, so the set of contributions and their coefficients are the
// reference's; only the execution (grouped, pull-style, on the device)
// differs.
#include "tjsq_plan.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <unordered_map>

namespace tjsq {

static constexpr int SU = 2;     // spin step, common.f90:11
static constexpr int NEIBT = 12; // common.f90:34

// --------------------------------------------------------------------------
// 6j symbol (W6JS, block.f90:1564-1615; FACTRL, block.f90:1508-1516)
// --------------------------------------------------------------------------
static const double* fact_table() {
  // initialised once, thread-safe (plans are compiled on several threads)
  static const std::array<double, 302> f = [] {
    std::array<double, 302> a{};
    a[0] = 1.0;
    for (int i = 1; i <= 142; ++i) a[i] = a[i - 1] * (double)(float)i;
    return a;
  }();
  return f.data();
}

double w6js(int j1, int j2, int j3, int l1, int l2, int l3) {
  const double* F = fact_table();
  const int ia = j1 + j2, ib = j1 - j2, ic = j1 + l2, id = j1 - l2;
  const int ie = l1 + j2, iff = l1 - j2, ig = l1 + l2, ih = l1 - l2;
  if (ia < j3 || std::abs(ib) > j3 || ic < l3 || std::abs(id) > l3 || ie < l3 || std::abs(iff) > l3 ||
      ig < j3 || std::abs(ih) > j3)
    return 0.0;
  const int s1 = ia + j3, s2 = ic + l3, s3 = ie + l3, s4 = ig + j3;
  if ((s1 | s2 | s3 | s4) & 1) return 0.0;
  const int wmin = std::max(std::max(s1, s2), std::max(s3, s4));
  const int ii = ia + ig, ij = j2 + j3 + l2 + l3, ik = j3 + j1 + l3 + l1;
  const int wmax = std::min(ii, std::min(ij, ik));
  double omega = 0.0;
  for (int w = wmin; w <= wmax; w += 2) {
    double den = F[(w - s1) / 2] * F[(w - s2) / 2] * F[(w - s3) / 2] * F[(w - s4) / 2] * F[(ii - w) / 2] *
                 F[(ij - w) / 2] * F[(ik - w) / 2];
    if (w % 4 != 0) den = -den;
    omega = omega + F[w / 2 + 1] / den;
  }
  const double t1 = F[(ia - j3) / 2] * F[(j3 + ib) / 2] * F[(j3 - ib) / 2] / F[s1 / 2 + 1];
  const double t2 = F[(ic - l3) / 2] * F[(l3 + id) / 2] * F[(l3 - id) / 2] / F[s2 / 2 + 1];
  const double t3 = F[(ie - l3) / 2] * F[(l3 + iff) / 2] * F[(l3 - iff) / 2] / F[s3 / 2 + 1];
  const double t4 = F[(ig - j3) / 2] * F[(j3 + ih) / 2] * F[(j3 - ih) / 2] / F[s4 / 2 + 1];
  return omega * std::sqrt(t1 * t2 * t3 * t4);
}

// 3j symbol (j1 j2 j3; m1 m2 m3), every argument twice the angular momentum
// (W3JS, block.f90:1522-1560): Racah's single sum over z,
//   (-1)^(j1-j2-m3) sqrt(D(j1 j2 j3) prod (j_i + m_i)! (j_i - m_i)!)
//   sum_z (-1)^z / [z! (j1+j2-j3-z)! (j1-m1-z)! (j2+m2-z)! (j3-j2+m1+z)! (j3-j1-m2+z)!]
// with D = (j1+j2-j3)! (j1-j2+j3)! (-j1+j2+j3)! / (j1+j2+j3+1)!; 0 where the
// selection rules fail
double w3js(int j1, int j2, int j3, int m1, int m2, int m3) {
  const double* F = fact_table();
  if (m1 + m2 + m3 != 0 || j3 > j1 + j2 || j3 < std::abs(j1 - j2)) return 0.0;
  if (((j1 + j2 + j3) | (j1 - m1) | (j2 - m2)) & 1) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
  const int a = j1 + j2 - j3, b = j1 - m1, c = j2 + m2, d = j3 - j2 + m1, e = j3 - j1 - m2;
  const int zmin = std::max(0, std::max(-d, -e)), zmax = std::min(a, std::min(b, c));
  double sum = 0.0;
  for (int z = zmin; z <= zmax; z += 2) {
    double den = F[z / 2] * F[(a - z) / 2] * F[(b - z) / 2] * F[(c - z) / 2] * F[(d + z) / 2] * F[(e + z) / 2];
    if (z % 4 != 0) den = -den;
    sum = sum + 1.0 / den;
  }
  const double tri = F[a / 2] * F[(j3 + j1 - j2) / 2] * F[(j3 - j1 + j2) / 2] / F[(j1 + j2 + j3 + 2) / 2];
  const double mm = F[(j1 + m1) / 2] * F[b / 2] * F[c / 2] * F[(j2 - m2) / 2] * F[(j3 - m3) / 2] * F[(j3 + m3) / 2];
  double v = sum * std::sqrt(tri * mm);
  if ((j1 - j2 - m3) % 4 != 0) v = -v;
  return v;
}

static inline double par(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

// --------------------------------------------------------------------------
// Fusion bases (Get_basis block.f90:9-104, QuickSort_Integer :212-239,
// Get_Unique_QN :165-198, Get_super_basis :110-161)
// --------------------------------------------------------------------------
int get_basis(int nb, const int* blk, int nsite, const int* site, std::vector<int>& sec, std::vector<int>& sub) {
  std::vector<std::pair<int, int>> pre;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nsite; ++j) {
      const int s1 = std::abs(blk[3 * i + 1] - site[3 * j + 1]), s2 = blk[3 * i + 1] + site[3 * j + 1];
      for (int ss = s1; ss <= s2; ss += SU) pre.emplace_back(blk[3 * i] + site[3 * j], ss);
    }
  const int n = (int)pre.size();
  for (int i = 0; i < n; ++i)  // 2S descending
    for (int x = i + 1; x < n; ++x)
      if (pre[i].second < pre[x].second) std::swap(pre[i], pre[x]);
  for (int i = 0; i < n; ++i)  // then N descending among equal 2S
    for (int x = i + 1; x < n; ++x)
      if (pre[i].second == pre[x].second && pre[i].first < pre[x].first) std::swap(pre[i], pre[x]);
  std::vector<std::pair<int, int>> uniq;
  {
    int start = 0;
    for (int i = 0; i < n; ++i) {
      int sublen = 0;
      for (int j = start; j < n; ++j)
        if (pre[j] == pre[i]) ++sublen;
      if (sublen) { start += sublen; uniq.push_back(pre[i]); }
    }
  }
  sec.clear();
  sub.clear();
  for (auto& q : uniq) {
    int dim = 0, cnt = 0;
    for (int x = 0; x < nb; ++x)
      for (int y = 0; y < nsite; ++y) {
        if (blk[3 * x] + site[3 * y] != q.first) continue;
        const int s1 = std::abs(blk[3 * x + 1] - site[3 * y + 1]), s2 = blk[3 * x + 1] + site[3 * y + 1];
        if (q.second >= s1 && q.second <= s2) {
          const int d = blk[3 * x + 2] * site[3 * y + 2];
          sub.insert(sub.end(), {blk[3 * x], blk[3 * x + 1], site[3 * y], site[3 * y + 1], dim, d});
          dim += d;
          ++cnt;
        }
      }
    sec.insert(sec.end(), {q.first, q.second, dim, cnt});
  }
  return (int)uniq.size();
}

int get_super_basis(const tjsq_basis& s, const tjsq_basis& e, int total_up, int total_down, std::vector<int>& wsec) {
  wsec.clear();
  for (int x = 0; x < s.nsec; ++x)
    for (int y = 0; y < e.nsec; ++y) {
      const int* a = s.sec + 4 * x;
      const int* b = e.sec + 4 * y;
      if (b[0] != total_up - a[0]) continue;
      const int s1 = std::abs(b[1] - a[1]), s2 = b[1] + a[1];
      if (total_down >= s1 && total_down <= s2) wsec.insert(wsec.end(), {a[0], a[1], b[0], b[1], a[2], b[2]});
    }
  return (int)(wsec.size() / 6);
}

// --------------------------------------------------------------------------
// Lattice tables (dmrg_setup0 dmrg_sub.f90:1-29, dmrg_setup :31-281 square,
// open = 1; couplings main.f90:35-47)
// --------------------------------------------------------------------------
static void square_neib(LatticeTables& L, int nleg, int nxc) {
  const int N = L.nsite;
  std::fill(L.neib.begin(), L.neib.end(), 0);
  std::fill(L.neiba.begin(), L.neiba.end(), 0);
  auto NB = [&](int i, int k) -> int& { return L.neib[(size_t)i * NEIBT + k - 1]; };
  auto NBA = [&](int i, int j) -> int& { return L.neiba[(size_t)i * (N + 1) + j]; };
  for (int ix = 1; ix <= nxc; ++ix)
    for (int iy = 1; iy <= nleg; ++iy) {
      const int ii = iy + (ix - 1) * nleg;
      int jy = ii - iy + (iy % nleg) + 1;
      if (iy != nleg || nleg != 2) { NB(ii, 1) = jy; NB(jy, 2) = ii; NBA(ii, jy) = 1; NBA(jy, ii) = 2; }
      int jx = ii + nleg;
      if (jx > nleg * nxc) jx = 0;
      if (jx) {
        NB(ii, 3) = jx; NB(jx, 4) = ii; NBA(ii, jx) = 3; NBA(jx, ii) = 4;
        jy = jx - 1 + (iy == 1 ? nleg : 0);
        if (jy) { NB(ii, 5) = jy; NB(jy, 6) = ii; NBA(ii, jy) = 5; NBA(jy, ii) = 6; }
        jy = jx + 1 - (iy == nleg ? nleg : 0);
        if (jy) { NB(ii, 7) = jy; NB(jy, 8) = ii; NBA(ii, jy) = 7; NBA(jy, ii) = 8; }
      }
    }
}

void lattice_setup(LatticeTables& L, int nleg, int ns, int ns21) {
  const int nn = ns + ns21 + 2, nxc = (nn - 1) / nleg + 1, totsite = nleg * nxc;
  std::fill(L.nposi.begin(), L.nposi.end(), 0);
  std::fill(L.nposi_lat.begin(), L.nposi_lat.end(), 0);
  std::fill(L.nposi_dm.begin(), L.nposi_dm.end(), 0);
  for (int i = 1; i <= ns + 1; ++i) { L.nposi[i] = 1; L.nposi_lat[i] = i; L.nposi_dm[i] = i; }
  L.nposi[ns + 1] = 2;
  for (int i = 1; i <= ns21 + 1; ++i) {
    L.nposi[ns + 1 + i] = 3;
    L.nposi_dm[totsite + 1 - i] = i + ns + 1;
    L.nposi_lat[i + ns + 1] = totsite + 1 - i;
  }
  L.nposi[ns + ns21 + 2] = 4;
  square_neib(L, nleg, nxc);
}

static void setup0(LatticeTables& L) {
  const int N = L.nsite;
  for (int ns = 1; ns <= N - 2; ++ns) {
    int k0 = 0;
    for (int i1 = 1; i1 <= ns; ++i1) {
      L.inb1[(size_t)i1 * (N + 1) + ns] = 10000;
      int k00 = 0;
      for (int j = 1; j <= NEIBT; ++j)
        if (L.neib_at(i1, j) >= ns) ++k00;
      if (k00) L.inb1[(size_t)i1 * (N + 1) + ns] = ++k0;
    }
    L.nleg11[ns] = k0;
  }
}

void build_lattice(const tjsq_lattice& lat, LatticeTables& L) {
  const int N = lat.ny * lat.nx_tables;
  L.nsite = N;
  L.neib.assign((size_t)(N + 1) * NEIBT, 0);
  L.neiba.assign((size_t)(N + 1) * (N + 1), 0);
  L.nposi.assign(N + 2, 0);
  L.nposi_lat.assign(N + 2, 0);
  L.nposi_dm.assign(N + 2, 0);
  L.inb1.assign((size_t)(N + 1) * (N + 1), 0);
  L.nleg11.assign(N + 1, 0);
  for (int k = 0; k <= NEIBT; ++k) L.jt[k] = L.jd[k] = L.jn[k] = 0.0;
  for (int k = 1; k <= 4; ++k) { L.jt[k] = lat.jt[0]; L.jd[k] = lat.jd[0]; L.jn[k] = lat.jn[0]; }
  for (int k = 5; k <= 8; ++k) { L.jt[k] = lat.jt[1]; L.jd[k] = lat.jd[1]; L.jn[k] = lat.jn[1]; }
  L.esite = lat.esite;
  setup0(L);
  lattice_setup(L, lat.ny, N / 2 - 1, N / 2 - 1);
  setup0(L);
}

// hmerge's integer part, get_sys.f90:4393-4405 and 4477-4511
void hmerge_tables(const LatticeTables& L, int sys_len, int env_len, Couplings& C) {
  std::memset(C.coup_bs, 0, sizeof(C.coup_bs));
  std::memset(C.coup_bb, 0, sizeof(C.coup_bb));
  C.inb.assign(L.nsite + 6, 0);
  const int dmsite = sys_len + env_len + 2, syssite = sys_len + 1;
  for (int k = 1; k <= NEIBT; ++k) {
    int j = L.neib_at(syssite, k);
    if (L.npos(j) == 3) C.coup_bs[2][1] = 1;
    if (L.npos(j) == 1) C.coup_bs[1][1] = 1;
    j = L.neib_at(L.nposi_lat[dmsite], k);
    if (L.npos(j) == 3) C.coup_bs[2][2] = 1;
    if (L.npos(j) == 1) C.coup_bs[1][2] = 1;
  }
  int k11 = 0;
  for (int I = 1; I <= sys_len; ++I) {
    int k1 = 0;
    for (int k = 1; k <= NEIBT; ++k)
      if (L.npos(L.neib_at(I, k)) == 3) ++k1;
    C.inb[I] = k1;
    if (L.nleg11[env_len] < L.nleg11[sys_len]) C.inb[I] = 0;
    if (C.inb[I]) ++k11;
  }
  C.coup_bb[1] = k11;
  C.inb[sys_len + 1] = 0;
  C.inb[sys_len + env_len + 2] = 0;
  for (int I = 1; I <= env_len; ++I) {
    int k1 = 0;
    for (int k = 1; k <= NEIBT; ++k) {
      const int J = L.neib_at(L.nposi_lat[I + sys_len + 1], k);
      if (L.npos(J) == 1 && C.inb[J] == 0) ++k1;
    }
    C.inb[I + sys_len + 1] = k1;
    if (k1) ++k11;
  }
  C.coup_bb[2] = k11 - C.coup_bb[1];
}

// --------------------------------------------------------------------------
// Contribution emission
// --------------------------------------------------------------------------
namespace {

// logical operand: stored matrix at off; logical = stored^T when stored_t
struct LOp {
  const Op* op;
  bool stored_t;   // transposed view (block_transfer_trans)
  std::vector<OpSub> vsub;  // logical sub-block metadata
};

LOp view(const Op& o) {
  LOp v{&o, false, o.sub};
  return v;
}

// block_transfer_trans, block.f90:1676-1706 (metadata only)
LOp trans_view(const Op& o) {
  LOp v{&o, true, {}};
  v.vsub.reserve(o.sub.size());
  for (const OpSub& s : o.sub) {
    OpSub t = s;
    t.num_up = s.num_up + o.up_dif;
    t.num_down = s.num_down + s.down_dif;
    t.down_dif = -s.down_dif;
    t.row_dim = s.sdim;
    t.sdim = s.row_dim;
    v.vsub.push_back(t);
  }
  return v;
}

int up_dif_of(const LOp& v) { return v.stored_t ? -v.op->up_dif : v.op->up_dif; }
int down_dif_of(const LOp& v) { return v.op->down_dif; }

int find_op(const LOp& v, int up, int down, bool chk, int dd) {
  for (size_t k = 0; k < v.vsub.size(); ++k)
    if (v.vsub[k].num_up == up && v.vsub[k].num_down == down && (!chk || v.vsub[k].down_dif == dd)) return (int)k;
  return -1;
}

struct Operand { int64_t off; int ld; uint8_t t; int rows, cols; };  // rows x cols of op(stored)

// op(logical sub s) with the BLAS flag `want_t`
Operand operand(const LOp& v, int s, bool want_t) {
  const OpSub& st = v.op->sub[s];  // stored metadata
  Operand r;
  r.off = st.off;
  r.ld = st.row_dim;                // stored leading dimension
  r.t = (uint8_t)(want_t != v.stored_t);
  r.rows = r.t ? st.sdim : st.row_dim;
  r.cols = r.t ? st.row_dim : st.sdim;
  return r;
}


const int ANYQ = -100000;

struct Emitter {
  HostPlan& P;
  // per wave sector: its sys/env basis sector indices and block-id base
  std::vector<int> wsys, wenv, wbase;
  std::vector<double> site_data;   // values of the site operators (1x1 blocks)

  bool bad_dims = false;             // an operator block disagrees with the basis dims
  explicit Emitter(HostPlan& p) : P(p) {}

  int find_sec(const BasisView& bs, int up, int down) const {
    for (size_t x = 0; x < bs.sec.size(); ++x)
      if (bs.sec[x].new_up == up && bs.sec[x].new_down == down) return (int)x;
    return -1;
  }
  int find_wave(int su, int sd, int eu, int ed) const {
    for (size_t k = 0; k < P.wsec.size(); ++k) {
      const WSec& w = P.wsec[k];
      if (w.sys_up == su && w.sys_down == sd && w.env_up == eu && w.env_down == ed) return (int)k;
    }
    return -1;
  }
  // enlarged-basis search: sector (up, down) then first sub with matching QNs
  int find_bs_sub(const BasisView& bs, int up, int down, int bl_up, int bl_down, int st_up, int st_down,
                  int* sec_out) const {
    for (size_t k = 0; k < bs.sec.size(); ++k) {
      if (bs.sec[k].new_up != up || bs.sec[k].new_down != down) continue;
      for (size_t l = 0; l < bs.sec[k].sub.size(); ++l) {
        const MapSub& m = bs.sec[k].sub[l];
        if (bl_up != ANYQ && m.bl_up != bl_up) continue;
        if (bl_down != ANYQ && m.bl_down != bl_down) continue;
        if (st_up != ANYQ && m.st_up != st_up) continue;
        if (st_down != ANYQ && m.st_down != st_down) continue;
        *sec_out = (int)k;
        return (int)l;
      }
    }
    return -1;
  }

  int block_id(int k, int ms, int me) const {
    const int ne = (int)P.env_bs.sec[wenv[k]].sub.size();
    return wbase[k] + ms * ne + me;
  }
  // input sub-block (wave sector i, sys sub ms, env sub me)
  void xblk(int i, int ms, int me, int64_t& off, int& ld) const {
    const WSec& w = P.wsec[i];
    const MapSub& s = P.sys_bs.sec[wsys[i]].sub[ms];
    const MapSub& e = P.env_bs.sec[wenv[i]].sub[me];
    ld = w.sys_dim;
    off = w.off + (int64_t)e.spos * w.sys_dim + s.spos;
  }

  void emit(uint8_t kind, int i, int ms, int me, int k, int ms2, int me2, const Operand* A, const Operand* B, int M,
            int N, int K1, int K2, double c) {
    if ((A && (A->rows != M || A->cols != K1)) || (B && (B->rows != K2 || B->cols != N))) {
      bad_dims = true;
      return;
    }
    // a vanishing coefficient (a zero 6j symbol or coupling) adds exactly
    // zero for finite operands: its product is not formed
    if (c == 0.0) return;
    Contrib q{};
    q.kind = kind;
    if (A) { q.a_off = A->off; q.a_ld = A->ld; q.a_t = A->t; }
    if (B) { q.b_off = B->off; q.b_ld = B->ld; q.b_t = B->t; }
    xblk(i, ms, me, q.x_off, q.x_ld);
    q.out_block = block_id(k, ms2, me2);
    q.M = M; q.N = N; q.K1 = K1; q.K2 = K2; q.c = c;
    P.contrib.push_back(q);
  }

  // ---- the 14 kernels (get_sys.f90) ----------------------------------------

  // block_vec_dia :1242-1289 (sys block scalar operator on sys sub-blocks)
  void block_vec_dia(const LOp& bl, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const int x = wsys[i];
      const MapSec& S = P.sys_bs.sec[x];
      for (size_t y = 0; y < S.sub.size(); ++y) {
        const MapSub& m = S.sub[y];
        const double c1 = coef / std::sqrt(1.0 + m.bl_down);
        const int k = find_op(bl, m.bl_up, m.bl_down, false, 0);
        if (k < 0) continue;
        Operand A = operand(bl, k, false);
        const MapSec& E = P.env_bs.sec[wenv[i]];
        for (size_t me = 0; me < E.sub.size(); ++me)
          emit(C_LEFT, (int)i, (int)y, (int)me, (int)i, (int)y, (int)me, &A, nullptr, m.sdim, E.sub[me].sdim, m.sdim,
               E.sub[me].sdim, c1);
      }
    }
  }

  // vec_block_dia :1293-1340
  void vec_block_dia(const LOp& bl, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t y = 0; y < E.sub.size(); ++y) {
        const MapSub& m = E.sub[y];
        const double c1 = coef / std::sqrt(1.0 + m.bl_down);
        const int k = find_op(bl, m.bl_up, m.bl_down, false, 0);
        if (k < 0) continue;
        Operand B = operand(bl, k, true);
        const MapSec& S = P.sys_bs.sec[wsys[i]];
        for (size_t ms = 0; ms < S.sub.size(); ++ms)
          emit(C_RIGHT, (int)i, (int)ms, (int)y, (int)i, (int)ms, (int)y, nullptr, &B, S.sub[ms].sdim, m.sdim,
               S.sub[ms].sdim, m.sdim, c1);
      }
    }
  }

  // block_site_vec_dia :1460-1530
  void block_site_vec_dia(const LOp& bl, const LOp& st, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t y = 0; y < S.sub.size(); ++y) {
        const MapSub& m = S.sub[y];
        const int b = find_op(bl, m.bl_up, m.bl_down, false, 0);
        const int s = find_op(st, m.st_up, m.st_down, false, 0);
        if (b < 0 || s < 0) continue;
        double c = coef * sval(st, s);
        c = c / std::sqrt((1.0 + m.bl_down) * (1.0 + m.st_down));
        Operand A = operand(bl, b, false);
        for (size_t me = 0; me < E.sub.size(); ++me)
          emit(C_LEFT, (int)i, (int)y, (int)me, (int)i, (int)y, (int)me, &A, nullptr, m.sdim, E.sub[me].sdim, m.sdim,
               E.sub[me].sdim, c);
      }
    }
  }

  // vec_block_site_dia :1534-1605
  void vec_block_site_dia(const LOp& bl, const LOp& st, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t y = 0; y < E.sub.size(); ++y) {
        const MapSub& m = E.sub[y];
        const int b = find_op(bl, m.bl_up, m.bl_down, false, 0);
        const int s = find_op(st, m.st_up, m.st_down, false, 0);
        if (b < 0 || s < 0) continue;
        double c = coef * sval(st, s);
        c = c / std::sqrt((1.0 + m.bl_down) * (1.0 + m.st_down));
        Operand B = operand(bl, b, true);
        for (size_t ms = 0; ms < S.sub.size(); ++ms)
          emit(C_RIGHT, (int)i, (int)ms, (int)y, (int)i, (int)ms, (int)y, nullptr, &B, S.sub[ms].sdim, m.sdim,
               S.sub[ms].sdim, m.sdim, c);
      }
    }
  }

  // site_vec_site_dia :1609-1705
  void site_vec_site_dia(const LOp& sst, const LOp& est, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t x = 0; x < S.sub.size(); ++x) {
        const MapSub& ms = S.sub[x];
        const int s = find_op(sst, ms.st_up, ms.st_down, false, 0);
        for (size_t y = 0; y < E.sub.size(); ++y) {
          const MapSub& me = E.sub[y];
          const int e = find_op(est, me.st_up, me.st_down, false, 0);
          if (s < 0 || e < 0) continue;
          double c = coef * sval(sst, s) * sval(est, e);
          c = c / std::sqrt((1.0 + ms.st_down) * (1.0 + me.st_down));
          emit(C_AXPY, (int)i, (int)x, (int)y, (int)i, (int)x, (int)y, nullptr, nullptr, ms.sdim, me.sdim, ms.sdim,
               me.sdim, c);
        }
      }
    }
  }

  // block_vec_site_dia :1709-1805
  void block_vec_site_dia(const LOp& sbl, const LOp& est, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t x = 0; x < S.sub.size(); ++x) {
        const MapSub& ms = S.sub[x];
        const int b = find_op(sbl, ms.bl_up, ms.bl_down, false, 0);
        for (size_t y = 0; y < E.sub.size(); ++y) {
          const MapSub& me = E.sub[y];
          const int e = find_op(est, me.st_up, me.st_down, false, 0);
          if (b < 0 || e < 0) continue;
          double c = coef * sval(est, e);
          c = c / std::sqrt((1.0 + me.st_down) * (1.0 + ms.bl_down));
          Operand A = operand(sbl, b, false);
          emit(C_LEFT, (int)i, (int)x, (int)y, (int)i, (int)x, (int)y, &A, nullptr, ms.sdim, me.sdim, ms.sdim, me.sdim,
               c);
        }
      }
    }
  }

  // site_vec_block_dia :1809-1906
  void site_vec_block_dia(const LOp& sst, const LOp& ebl, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t x = 0; x < S.sub.size(); ++x) {
        const MapSub& ms = S.sub[x];
        const int s = find_op(sst, ms.st_up, ms.st_down, false, 0);
        for (size_t y = 0; y < E.sub.size(); ++y) {
          const MapSub& me = E.sub[y];
          const int e = find_op(ebl, me.bl_up, me.bl_down, false, 0);
          if (s < 0 || e < 0) continue;
          double c = coef * sval(sst, s);
          c = c / std::sqrt((1.0 + me.bl_down) * (1.0 + ms.st_down));
          Operand B = operand(ebl, e, true);
          emit(C_RIGHT, (int)i, (int)x, (int)y, (int)i, (int)x, (int)y, nullptr, &B, ms.sdim, me.sdim, ms.sdim, me.sdim,
               c);
        }
      }
    }
  }

  // block_vec_block_dia :1910-2012
  void block_vec_block_dia(const LOp& sbl, const LOp& ebl, double coef) {
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t x = 0; x < S.sub.size(); ++x) {
        const MapSub& ms = S.sub[x];
        const int b = find_op(sbl, ms.bl_up, ms.bl_down, false, 0);
        for (size_t y = 0; y < E.sub.size(); ++y) {
          const MapSub& me = E.sub[y];
          const int e = find_op(ebl, me.bl_up, me.bl_down, false, 0);
          if (b < 0 || e < 0) continue;
          const double c1 = coef / std::sqrt((1.0 + me.bl_down) * (1.0 + ms.bl_down));
          Operand A = operand(sbl, b, false), B = operand(ebl, e, true);
          emit(C_BOTH, (int)i, (int)x, (int)y, (int)i, (int)x, (int)y, &A, &B, ms.sdim, me.sdim, ms.sdim, me.sdim, c1);
        }
      }
    }
  }

  // block_site_vec_ndia :2016-2265 (live :2049-2191)
  void block_site_vec_ndia(const LOp& bl, char fbl, const LOp& st, char fst, double coef, std::string& err) {
    (void)fbl;
    const int up_dif = up_dif_of(bl), d1 = down_dif_of(bl);
    if (up_dif != up_dif_of(st) || d1 != down_dif_of(st)) { err = "QN error in block_site_vec"; return; }
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const WSec& w = P.wsec[i];
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t x = 0; x < S.sub.size(); ++x) {
        const MapSub& mo = S.sub[x];
        for (int dd = -d1; dd <= d1; dd += SU)
          for (int sdd = -d1; sdd <= d1; sdd += SU) {
            const int b = find_op(bl, mo.bl_up, mo.bl_down, true, dd);
            const int nst_up = mo.st_up - up_dif, nst_down = mo.st_down - sdd;
            const int s = find_op(st, nst_up, nst_down, true, sdd);
            const int nbl_up = mo.bl_up + up_dif, nbl_down = mo.bl_down + dd;
            if (b < 0 || s < 0) continue;
            int y = -1;
            for (size_t q = 0; q < S.sub.size(); ++q) {
              const MapSub& mq = S.sub[q];
              if (mq.bl_up == nbl_up && mq.bl_down == nbl_down && mq.st_up == nst_up && mq.st_down == nst_down) {
                y = (int)q;
                break;
              }
            }
            if (y < 0) continue;
            const MapSub& mn = S.sub[y];
            const double sign_st = (fst == 'B') ? 1.0 : par(mo.bl_down);
            const int j1 = nbl_down, j2 = mo.bl_down, j3 = d1, j4 = mo.st_down, j5 = nst_down, j6 = w.sys_down;
            double c1 = w6js(j1, j2, j3, j4, j5, j6);
            if (c1 == 0.0) continue;
            c1 = c1 * par((j2 + j5 + j3 + j6) / 2) / std::sqrt(1.0 + d1);
            const double ct = (1.0 * sign_st * c1) * coef * sval(st, s);
            Operand A = operand(bl, b, false), AT = operand(bl, b, true);
            for (size_t me = 0; me < E.sub.size(); ++me) {
              const int nn = E.sub[me].sdim;
              emit(C_LEFT, (int)i, (int)x, (int)me, (int)i, y, (int)me, &A, nullptr, mn.sdim, nn, mo.sdim, nn, ct);
              if (d1 == 1)
                emit(C_LEFT, (int)i, y, (int)me, (int)i, (int)x, (int)me, &AT, nullptr, mo.sdim, nn, mn.sdim, nn, ct);
            }
          }
      }
    }
  }

  // vec_block_site_ndia :2269-2552 (live :2303-2440)
  void vec_block_site_ndia(const LOp& bl, char fbl, const LOp& st, char fst, double coef, std::string& err) {
    const int up_dif = up_dif_of(bl), d1 = down_dif_of(bl);
    if (up_dif != up_dif_of(st) || d1 != down_dif_of(st)) { err = "QN error in vec_block_site"; return; }
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const WSec& w = P.wsec[i];
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t x = 0; x < E.sub.size(); ++x) {
        const MapSub& mo = E.sub[x];
        for (int bdd = -d1; bdd <= d1; bdd += SU) {
          const int b = find_op(bl, mo.bl_up, mo.bl_down, true, bdd);
          for (int sdd = -d1; sdd <= d1; sdd += SU) {
            const int nst_up = mo.st_up - up_dif, nst_down = mo.st_down - sdd;
            const int s = find_op(st, nst_up, nst_down, true, sdd);
            const int nbl_up = mo.bl_up + up_dif, nbl_down = mo.bl_down + bdd;
            if (b < 0 || s < 0) continue;
            int y = -1;
            for (size_t q = 0; q < E.sub.size(); ++q) {
              const MapSub& mq = E.sub[q];
              if (mq.bl_up == nbl_up && mq.bl_down == nbl_down && mq.st_up == nst_up && mq.st_down == nst_down) {
                y = (int)q;
                break;
              }
            }
            if (y < 0) continue;
            const MapSub& mn = E.sub[y];
            const double sign_bl = (fbl == 'B') ? 1.0 : par(w.sys_down + nst_down);
            const double sign_st = (fst == 'B') ? 1.0 : par(w.sys_down);
            const int j1 = nbl_down, j2 = mo.bl_down, j3 = d1, j4 = mo.st_down, j5 = nst_down, j6 = w.env_down;
            double c1 = w6js(j1, j2, j3, j4, j5, j6);
            if (c1 == 0.0) continue;
            c1 = c1 * par((j2 + j5 + j3 + j6) / 2) / std::sqrt(1.0 + d1);
            const double ct = (sign_bl * sign_st * c1) * coef * sval(st, s);
            Operand BT = operand(bl, b, true), B = operand(bl, b, false);
            for (size_t ms = 0; ms < S.sub.size(); ++ms) {
              const int mm = S.sub[ms].sdim;
              emit(C_RIGHT, (int)i, (int)ms, (int)x, (int)i, (int)ms, y, nullptr, &BT, mm, mn.sdim, mm, mo.sdim, ct);
              if (d1 == 1)
                emit(C_RIGHT, (int)i, (int)ms, y, (int)i, (int)ms, (int)x, nullptr, &B, mm, mo.sdim, mm, mn.sdim, ct);
            }
          }
        }
      }
    }
  }

  // block_vec_block_ndia :2555-2981 (live :2590-2834); ebl is the transposed
  // env operator of model_wave (block_transfer_trans)
  void block_vec_block_ndia(const LOp& sbl, char fsl, const LOp& ebl, char fel, double coef, std::string& err) {
    (void)fsl;
    const int up_dif = up_dif_of(sbl), d1 = down_dif_of(sbl);
    if (up_dif != up_dif_of(ebl) || d1 != down_dif_of(ebl)) { err = "QN error in block_vec_block"; return; }
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const WSec& w = P.wsec[i];
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t m = 0; m < S.sub.size(); ++m)
        for (size_t n = 0; n < E.sub.size(); ++n) {
          const MapSub& ms = S.sub[m];
          const MapSub& me = E.sub[n];
          for (int dd = -d1; dd <= d1; dd += SU) {
            const int b = find_op(sbl, ms.bl_up, ms.bl_down, true, dd);
            const int nsb_up = ms.bl_up + up_dif, nsb_down = ms.bl_down + dd;
            for (int edd = -d1; edd <= d1; edd += SU) {
              const int neb_up = me.bl_up - up_dif, neb_down = me.bl_down - edd;
              const int e = find_op(ebl, neb_up, neb_down, true, edd);
              if (b < 0 || e < 0) continue;
              for (int nsd = std::abs(w.sys_down - d1); nsd <= w.sys_down + d1; nsd += SU) {
                const int ns_up = w.sys_up + up_dif, ne_up = w.env_up - up_dif, ned = nsd;
                int ks = -1, ke = -1;
                const int ys = find_bs_sub(P.sys_bs, ns_up, nsd, nsb_up, nsb_down, ms.st_up, ms.st_down, &ks);
                if (ys < 0) continue;
                const int ye = find_bs_sub(P.env_bs, ne_up, ned, neb_up, neb_down, me.st_up, me.st_down, &ke);
                if (ye < 0) continue;
                const int k = find_wave(ns_up, nsd, ne_up, ned);
                if (k < 0) continue;
                const MapSub& mns = P.sys_bs.sec[ks].sub[ys];
                const MapSub& mne = P.env_bs.sec[ke].sub[ye];
                const double sign_el = (fel == 'B') ? 1.0 : par(w.sys_down + me.st_down);
                const int j1 = nsd, j2 = d1, j3 = w.sys_down, j4 = ms.bl_down, j5 = ms.st_down, j6 = nsb_down;
                const int j11 = ned, j22 = d1, j33 = w.env_down, j44 = me.bl_down, j55 = me.st_down, j66 = neb_down;
                double c1 = w6js(j1, j2, j3, j4, j5, j6);
                if (c1 == 0.0) continue;
                const double c11 = w6js(j11, j22, j33, j44, j55, j66);
                if (c11 == 0.0) continue;
                c1 = c1 * c11;
                if (c1 == 0.0) continue;
                c1 = c1 * std::sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + d1));
                if (((j6 + j5 + j3 + j2) / 2 + (j66 + j55 + j33 + j22) / 2) % 2 == 1) c1 = -c1;
                const double ct = (1.0 * sign_el) * coef * c1;
                Operand A = operand(sbl, b, false), B = operand(ebl, e, false);
                emit(C_BOTH, (int)i, (int)m, (int)n, k, ys, ye, &A, &B, mns.sdim, mne.sdim, ms.sdim, me.sdim, ct);
                if (d1 == 1) {
                  Operand AT = operand(sbl, b, true), BT = operand(ebl, e, true);
                  emit(C_BOTH, k, ys, ye, (int)i, (int)m, (int)n, &AT, &BT, ms.sdim, me.sdim, mns.sdim, mne.sdim, ct);
                }
              }
            }
          }
        }
    }
  }

  // block_vec_site_ndia :2984-3351 (live :3020-3212)
  void block_vec_site_ndia(const LOp& sbl, char fsl, const LOp& est, char fet, double coef, std::string& err) {
    (void)fsl;
    const int up_dif = up_dif_of(sbl), d1 = down_dif_of(sbl);
    if (up_dif != up_dif_of(est) || d1 != down_dif_of(est)) { err = "QN error in block_vec_site"; return; }
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const WSec& w = P.wsec[i];
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t m = 0; m < S.sub.size(); ++m)
        for (size_t n = 0; n < E.sub.size(); ++n) {
          const MapSub& ms = S.sub[m];
          const MapSub& me = E.sub[n];
          for (int bdd = -d1; bdd <= d1; bdd += SU) {
            const int nsb_up = ms.bl_up + up_dif, nsb_down = ms.bl_down + bdd;
            const int b = find_op(sbl, ms.bl_up, ms.bl_down, true, bdd);
            for (int sdd = -d1; sdd <= d1; sdd += SU) {
              const int nes_up = me.st_up - up_dif, nes_down = me.st_down - sdd;
              const int s = find_op(est, nes_up, nes_down, true, sdd);
              const int ns_up = w.sys_up + up_dif, ne_up = w.env_up - up_dif;
              if (b < 0 || s < 0) continue;
              for (int nsd = std::abs(w.sys_down - d1); nsd <= w.sys_down + d1; nsd += SU) {
                const int ned = nsd;
                int ks = -1, ke = -1;
                const int ys = find_bs_sub(P.sys_bs, ns_up, nsd, nsb_up, nsb_down, ANYQ, ms.st_down, &ks);
                if (ys < 0) continue;
                const int ye = find_bs_sub(P.env_bs, ne_up, ned, ANYQ, me.bl_down, nes_up, nes_down, &ke);
                if (ye < 0) continue;
                const int k = find_wave(ns_up, nsd, ne_up, ned);
                if (k < 0) continue;
                const MapSub& mns = P.sys_bs.sec[ks].sub[ys];
                const double sign_et = (fet == 'B') ? 1.0 : par(w.sys_down);
                const int j1 = nsd, j2 = d1, j3 = w.sys_down, j4 = ms.bl_down, j5 = ms.st_down, j6 = nsb_down;
                const int j11 = ned, j22 = d1, j33 = w.env_down, j44 = me.st_down, j55 = me.bl_down, j66 = nes_down;
                double c1 = w6js(j1, j2, j3, j4, j5, j6);
                if (c1 == 0.0) continue;
                const double c11 = w6js(j11, j22, j33, j44, j55, j66);
                if (c11 == 0.0) continue;
                c1 = c1 * c11;
                if (c1 == 0.0) continue;
                c1 = c1 * std::sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + d1));
                if (((j6 + j5 + j3 + j2) / 2 + (j55 + j44 + j11 + j22) / 2) % 2 != 0) c1 = -c1;
                const double ct = (1.0 * sign_et * c1) * coef * sval(est, s);
                Operand A = operand(sbl, b, false), AT = operand(sbl, b, true);
                emit(C_LEFT, (int)i, (int)m, (int)n, k, ys, ye, &A, nullptr, mns.sdim, me.sdim, ms.sdim, me.sdim, ct);
                if (d1 == 1)
                  emit(C_LEFT, k, ys, ye, (int)i, (int)m, (int)n, &AT, nullptr, ms.sdim, me.sdim, mns.sdim, me.sdim, ct);
              }
            }
          }
        }
    }
  }

  // site_vec_block_ndia :3355-3757 (live :3390-3605); ebl transposed view
  void site_vec_block_ndia(const LOp& sst, char fst, const LOp& ebl, char fel, double coef, std::string& err) {
    const int up_dif = up_dif_of(sst), d1 = down_dif_of(sst);
    if (up_dif != up_dif_of(ebl) || d1 != down_dif_of(ebl)) { err = "QN error in site_vec_block"; return; }
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const WSec& w = P.wsec[i];
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t m = 0; m < S.sub.size(); ++m)
        for (size_t n = 0; n < E.sub.size(); ++n) {
          const MapSub& ms = S.sub[m];
          const MapSub& me = E.sub[n];
          for (int sdd = -d1; sdd <= d1; sdd += SU) {
            const int nss_up = ms.st_up + up_dif, nss_down = ms.st_down + sdd;
            const int s = find_op(sst, ms.st_up, ms.st_down, true, sdd);
            for (int bdd = -d1; bdd <= d1; bdd += SU) {
              const int neb_up = me.bl_up - up_dif, neb_down = me.bl_down - bdd;
              const int e = find_op(ebl, neb_up, neb_down, true, bdd);
              const int ns_up = w.sys_up + up_dif, ne_up = w.env_up - up_dif;
              if (s < 0 || e < 0) continue;
              for (int nsd = std::abs(w.sys_down - d1); nsd <= w.sys_down + d1; nsd += SU) {
                const int ned = nsd;
                int ks = -1, ke = -1;
                const int ys = find_bs_sub(P.sys_bs, ns_up, nsd, ms.bl_up, ms.bl_down, nss_up, nss_down, &ks);
                if (ys < 0) continue;
                const int ye = find_bs_sub(P.env_bs, ne_up, ned, neb_up, neb_down, me.st_up, me.st_down, &ke);
                if (ye < 0) continue;
                const int k = find_wave(ns_up, nsd, ne_up, ned);
                if (k < 0) continue;
                const MapSub& mne = P.env_bs.sec[ke].sub[ye];
                const int j1 = nsd, j2 = d1, j3 = w.sys_down, j4 = ms.st_down, j5 = ms.bl_down, j6 = nss_down;
                const int j11 = ned, j22 = d1, j33 = w.env_down, j44 = me.bl_down, j55 = me.st_down, j66 = neb_down;
                double c1 = w6js(j1, j2, j3, j4, j5, j6);
                if (c1 == 0.0) continue;
                const double c11 = w6js(j11, j22, j33, j44, j55, j66);
                if (c11 == 0.0) continue;
                c1 = c1 * c11;
                if (c1 == 0.0) continue;
                c1 = c1 * std::sqrt((1.0 + j11) * (j33 + 1.0) / (1.0 + d1));
                if (((j66 + j55 + j33 + j22) / 2 + (j5 + j4 + j1 + j2) / 2) % 2 != 0) c1 = -c1;
                const double sign_st = (fst == 'B') ? 1.0 : par(ms.bl_down);
                const double sign_el = (fel == 'B') ? 1.0 : par(w.sys_down + me.st_down);
                const double ct = (sign_st * sign_el * c1) * coef * sval(sst, s);
                Operand B = operand(ebl, e, false), BT = operand(ebl, e, true);
                emit(C_RIGHT, (int)i, (int)m, (int)n, k, ys, ye, nullptr, &B, ms.sdim, mne.sdim, ms.sdim, me.sdim, ct);
                if (d1 == 1)
                  emit(C_RIGHT, k, ys, ye, (int)i, (int)m, (int)n, nullptr, &BT, ms.sdim, me.sdim, ms.sdim, mne.sdim,
                       ct);
              }
            }
          }
        }
    }
  }

  // site_vec_site_ndia :3761-4124 (live :3794-3994)
  void site_vec_site_ndia(const LOp& sst, char fst, const LOp& est, char fet, double coef, std::string& err) {
    const int up_dif = up_dif_of(sst), d1 = down_dif_of(sst);
    if (up_dif != up_dif_of(est) || d1 != down_dif_of(est)) { err = "QN error in site_vec_site"; return; }
    for (size_t i = 0; i < P.wsec.size(); ++i) {
      const WSec& w = P.wsec[i];
      const MapSec& S = P.sys_bs.sec[wsys[i]];
      const MapSec& E = P.env_bs.sec[wenv[i]];
      for (size_t m = 0; m < S.sub.size(); ++m)
        for (size_t n = 0; n < E.sub.size(); ++n) {
          const MapSub& ms = S.sub[m];
          const MapSub& me = E.sub[n];
          for (int sd1 = -d1; sd1 <= d1; sd1 += SU) {
            const int s = find_op(sst, ms.st_up, ms.st_down, true, sd1);
            const int nss_up = ms.st_up + up_dif, nss_down = ms.st_down + sd1;
            for (int sd2 = -d1; sd2 <= d1; sd2 += SU) {
              const int nes_up = me.st_up - up_dif, nes_down = me.st_down - sd2;
              const int e = find_op(est, nes_up, nes_down, true, sd2);
              const int ns_up = w.sys_up + up_dif, ne_up = w.env_up - up_dif;
              if (s < 0 || e < 0) continue;
              for (int nsd = std::abs(w.sys_down - d1); nsd <= w.sys_down + d1; nsd += SU) {
                const int ned = nsd;
                int ks = -1, ke = -1;
                const int ys = find_bs_sub(P.sys_bs, ns_up, nsd, ANYQ, ms.bl_down, nss_up, nss_down, &ks);
                if (ys < 0) continue;
                const int ye = find_bs_sub(P.env_bs, ne_up, ned, ANYQ, me.bl_down, nes_up, nes_down, &ke);
                if (ye < 0) continue;
                const int k = find_wave(ns_up, nsd, ne_up, ned);
                if (k < 0) continue;
                const double sign_st = (fst == 'B') ? 1.0 : par(ms.bl_down);
                const double sign_et = (fet == 'B') ? 1.0 : par(w.sys_down);
                const double ct = (sign_st * sign_et) * coef;
                const int j1 = nsd, j2 = d1, j3 = w.sys_down, j4 = ms.st_down, j5 = ms.bl_down, j6 = nss_down;
                const int j11 = ned, j22 = d1, j33 = w.env_down, j44 = me.st_down, j55 = me.bl_down, j66 = nes_down;
                double c1 = w6js(j1, j2, j3, j4, j5, j6);
                if (c1 == 0.0) continue;
                const double c11 = w6js(j11, j22, j33, j44, j55, j66);
                if (c11 == 0.0) continue;
                c1 = c1 * c11;
                if (c1 == 0.0) continue;
                c1 = c1 * std::sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + d1));
                if (((j55 + j44 + j11 + j22) / 2 + (j5 + j4 + j1 + j2) / 2) % 2 != 0) c1 = -c1;
                c1 = ct * c1 * sval(sst, s) * sval(est, e);
                emit(C_AXPY, (int)i, (int)m, (int)n, k, ys, ye, nullptr, nullptr, ms.sdim, me.sdim, ms.sdim, me.sdim, c1);
                if (d1 == 1)
                  emit(C_AXPY, k, ys, ye, (int)i, (int)m, (int)n, nullptr, nullptr, ms.sdim, me.sdim, ms.sdim, me.sdim,
                       c1);
              }
            }
          }
        }
    }
  }

  // 1x1 site operator value: site operators live in the host-only site arena
  double sval(const LOp& v, int s) const { return site_data[(size_t)v.op->sub[s].off]; }
};

// Site operators (Get_site_operator_su2_U0, dmrg_sub.f90:370-473) with the
// particle label ub; their 1x1 values live in Emitter::site_data.
struct SiteOps {
  Op elec_up, num_elec, st_sd, st_si;
  std::vector<double> vals;
};
SiteOps make_site_ops(int ub) {
  SiteOps s;
  s.vals = {-std::sqrt(2.0), 0.0, std::sqrt(2.0), std::sqrt(1.5), 1.0, std::sqrt(2.0)};
  s.elec_up.len = 1; s.elec_up.up_dif = ub; s.elec_up.down_dif = 1;
  s.elec_up.sub = {{0, 0, 1, 1, 1, 0}};
  s.num_elec.len = 1; s.num_elec.up_dif = 0; s.num_elec.down_dif = 0;
  s.num_elec.sub = {{0, 0, 0, 1, 1, 1}, {ub, 1, 0, 1, 1, 2}};
  s.st_sd.len = 1; s.st_sd.up_dif = 0; s.st_sd.down_dif = SU;
  s.st_sd.sub = {{ub, 1, 0, 1, 1, 3}};
  s.st_si.len = 1; s.st_si.up_dif = 0; s.st_si.down_dif = 0;
  s.st_si.sub = {{0, 0, 0, 1, 1, 4}, {ub, 1, 0, 1, 1, 5}};
  return s;
}

// block_pass_info (block.f90:1741-1766): same structure, fresh zeroed storage
Op pass_info(const Op& t, int64_t& next) {
  Op o = t;
  for (OpSub& s : o.sub) {
    s.off = next;
    next += (int64_t)s.row_dim * s.sdim;
  }
  return o;
}

// block_add_block_two (block.f90:1782-1814) as merge recipes
void add_block_two(const Op& b1, double coef, const Op& nb, std::vector<MergeTerm>& out) {
  size_t x1 = 0;
  for (const OpSub& d : nb.sub) {
    for (size_t x = x1; x < b1.sub.size(); ++x) {
      const OpSub& s = b1.sub[x];
      if (s.num_up == d.num_up && s.num_down == d.num_down && s.down_dif == d.down_dif) {
        out.push_back({d.off, s.off, (int64_t)d.row_dim * d.sdim, coef});
        x1 = x + 1;
        break;
      }
    }
  }
}

struct MSet { Op elec_up, elec_down, spin_sd, num_sn; bool has = false; };

int read_opset(const tjsq_opset& in, int64_t base, OpSetView& out, std::string& msg) {
  if (in.nops < 1 || (in.nops - 1) % 4 != 0) { msg = "operator count must be 1 + 4 * active sites"; return TJSQ_ERR_ARG; }
  int first = 0;
  auto read = [&](int op, Op& o) -> int {
    const int* h = in.hdr + 4 * op;
    o.len = h[0]; o.up_dif = h[1]; o.down_dif = h[2];
    o.sub.resize(h[3]);
    for (int s = 0; s < h[3]; ++s) {
      const int* q = in.sub + 5 * (first + s);
      OpSub& d = o.sub[s];
      d.num_up = q[0]; d.num_down = q[1]; d.down_dif = q[2]; d.row_dim = q[3]; d.sdim = q[4];
      d.off = base + in.off[first + s];
      if (d.row_dim < 0 || d.sdim < 0 || in.off[first + s] < 0 ||
          in.off[first + s] + (int64_t)d.row_dim * d.sdim > in.ndata) {
        msg = "operator sub-block outside the data array";
        return TJSQ_ERR_ARG;
      }
    }
    first += h[3];
    return TJSQ_OK;
  };
  int rc = read(0, out.ham);
  if (rc) return rc;
  out.site.resize(in.nops - 1);
  for (int op = 1; op < in.nops; ++op)
    if ((rc = read(op, out.site[op - 1]))) return rc;
  return TJSQ_OK;
}

}  // namespace

int build_host_plan(const tjsq_lattice& lat, int sys_len, int env_len, const tjsq_opset& sys, const tjsq_opset& env,
                    const tjsq_basis& sys_bs, const tjsq_basis& env_bs, int nw, const int* wsec, const Shape& shape,
                    HostPlan& P, std::string& msg) {
  P = HostPlan();
  P.sys_len = sys_len;
  P.env_len = env_len;
  if (lat.ny < 1 || lat.nx_tables < 1 || sys_len < 1 || env_len < 1 || sys_len + env_len + 2 > lat.ny * lat.nx_tables) {
    msg = "lattice too small for sys_len + env_len + 2 sites";
    return TJSQ_ERR_ARG;
  }
  P.sys_base = 0;
  P.env_base = sys.ndata;
  int rc = read_opset(sys, P.sys_base, P.sys, msg);
  if (rc) return rc;
  if ((rc = read_opset(env, P.env_base, P.env, msg))) return rc;
  auto read_basis = [](const tjsq_basis& b, BasisView& v) {
    v.sec.resize(b.nsec);
    const int* q = b.sub;
    for (int k = 0; k < b.nsec; ++k) {
      MapSec& s = v.sec[k];
      s.new_up = b.sec[4 * k]; s.new_down = b.sec[4 * k + 1]; s.dim = b.sec[4 * k + 2];
      s.sub.resize(b.sec[4 * k + 3]);
      for (auto& m : s.sub) { m = {q[0], q[1], q[2], q[3], q[4], q[5]}; q += 6; }
    }
  };
  read_basis(sys_bs, P.sys_bs);
  read_basis(env_bs, P.env_bs);
  P.wsec.resize(nw);
  int64_t off = 0;
  for (int i = 0; i < nw; ++i) {
    const int* w = wsec + 6 * i;
    P.wsec[i] = {w[0], w[1], w[2], w[3], w[4], w[5], off};
    off += (int64_t)w[4] * w[5];
  }
  P.dim = off;

  Emitter em(P);
  em.wsys.resize(nw);
  em.wenv.resize(nw);
  em.wbase.resize(nw);
  int nblk = 0, ngroup = 0;
  for (int i = 0; i < nw; ++i) {
    const WSec& w = P.wsec[i];
    const int xs = em.find_sec(P.sys_bs, w.sys_up, w.sys_down), xe = em.find_sec(P.env_bs, w.env_up, w.env_down);
    if (xs < 0 || xe < 0) { msg = "wavefunction sector missing from an enlarged basis"; return TJSQ_ERR_ARG; }
    if (P.sys_bs.sec[xs].dim != w.sys_dim || P.env_bs.sec[xe].dim != w.env_dim) {
      msg = "wavefunction sector dims disagree with the bases";
      return TJSQ_ERR_ARG;
    }
    em.wsys[i] = xs;
    em.wenv[i] = xe;
    em.wbase[i] = nblk;
    const MapSec& S = P.sys_bs.sec[xs];
    const MapSec& E = P.env_bs.sec[xe];
    for (const MapSub& ms : S.sub)
      for (size_t e = 0; e < E.sub.size(); ++e) {
        const MapSub& me = E.sub[e];
        P.out_blocks.push_back(
            {w.off + (int64_t)me.spos * w.sys_dim + ms.spos, w.sys_dim, ms.sdim, me.sdim, ngroup + (int)e});
      }
    ngroup += (int)E.sub.size();
    nblk += (int)(S.sub.size() * E.sub.size());
  }

  // lattice + hmerge
  LatticeTables L;
  build_lattice(lat, L);
  lattice_setup(L, lat.ny, sys_len, env_len);
  Couplings C;
  hmerge_tables(L, sys_len, env_len, C);

  SiteOps so = make_site_ops(lat.up_bias);
  em.site_data = so.vals;
  const int nact_s = (int)P.sys.site.size() / 4, nact_e = (int)P.env.site.size() / 4;
  auto need = [&](int idx, int nact, const char* side) -> bool {
    if (idx < 1 || idx > nact) {
      msg = std::string("active-site index out of range on the ") + side + " block";
      return false;
    }
    return true;
  };

  // merged operators (hmerge, get_sys.f90:4407-4475 and 4513-4591)
  int64_t next = P.env_base + env.ndata;
  P.merged_base = next;
  MSet ssnc[3], esnc[3];
  if (nact_s < 1 || nact_e < 1) { msg = "blocks need at least one active site"; return TJSQ_ERR_ARG; }
  for (int ia = 1; ia <= 2; ++ia) {
    if (C.coup_bs[1][ia] == 1) {
      ssnc[ia].elec_up = pass_info(P.sys.op(1, 0), next);
      ssnc[ia].spin_sd = pass_info(P.sys.op(1, 2), next);
      ssnc[ia].num_sn = pass_info(P.sys.op(1, 3), next);
      ssnc[ia].has = true;
    }
    if (C.coup_bs[2][ia] == 1) {
      esnc[ia].elec_up = pass_info(P.env.op(1, 0), next);
      esnc[ia].elec_down = pass_info(P.env.op(1, 1), next);
      esnc[ia].spin_sd = pass_info(P.env.op(1, 2), next);
      esnc[ia].num_sn = pass_info(P.env.op(1, 3), next);
      esnc[ia].has = true;
    }
  }
  for (int ia = 1; ia <= 2; ++ia) {
    const int j = sys_len + 1 + (env_len + 1) * (ia - 1);
    const int sj = L.nposi_lat[j];
    for (int k1 = 1; k1 <= NEIBT; ++k1) {
      const int Ib = L.neib_at(sj, k1);
      const double jt1 = L.jt[k1], jd1 = L.jd[k1], jn1 = L.jn[k1];
      if (L.npos(Ib) == 3 && C.coup_bs[2][ia] != 0) {
        const int ib1 = L.inb1_at(L.nposi_dm[Ib] - sys_len - 1, env_len);
        if (!need(ib1, nact_e, "env")) return TJSQ_ERR_ARG;
        if (jt1 != 0.0) {
          add_block_two(P.env.op(ib1, 0), jt1, esnc[ia].elec_up, P.merge);
          add_block_two(P.env.op(ib1, 1), jt1, esnc[ia].elec_down, P.merge);
        }
        if (jd1 != 0.0) add_block_two(P.env.op(ib1, 2), jd1, esnc[ia].spin_sd, P.merge);
        if (jn1 != 0.0) add_block_two(P.env.op(ib1, 3), jn1, esnc[ia].num_sn, P.merge);
      }
      if (L.npos(Ib) == 1 && C.coup_bs[1][ia] == 1) {
        const int ib1 = L.inb1_at(L.nposi_dm[Ib], sys_len);
        if (!need(ib1, nact_s, "sys")) return TJSQ_ERR_ARG;
        if (jt1 != 0.0) add_block_two(P.sys.op(ib1, 0), jt1, ssnc[ia].elec_up, P.merge);
        if (jd1 != 0.0) add_block_two(P.sys.op(ib1, 2), jd1, ssnc[ia].spin_sd, P.merge);
        if (jn1 != 0.0) add_block_two(P.sys.op(ib1, 3), jn1, ssnc[ia].num_sn, P.merge);
      }
    }
  }
  std::vector<MSet> esub(C.coup_bb[1] + 1), ssub(C.coup_bb[2] + 1);
  for (int i = 1; i <= C.coup_bb[1]; ++i) {
    esub[i].elec_up = pass_info(P.env.op(1, 0), next);
    esub[i].elec_down = pass_info(P.env.op(1, 1), next);
    esub[i].spin_sd = pass_info(P.env.op(1, 2), next);
    esub[i].num_sn = pass_info(P.env.op(1, 3), next);
  }
  for (int i = 1; i <= C.coup_bb[2]; ++i) {
    ssub[i].elec_up = pass_info(P.sys.op(1, 0), next);
    ssub[i].spin_sd = pass_info(P.sys.op(1, 2), next);
    ssub[i].num_sn = pass_info(P.sys.op(1, 3), next);
  }
  {
    int jc = 0;
    for (int I = 1; I <= sys_len + env_len + 1; ++I) {
      if (C.inb[I] == 0) continue;
      const int Ia = L.nposi_lat[I];
      ++jc;
      for (int k1 = 1; k1 <= NEIBT; ++k1) {
        const int Ib = L.neib_at(Ia, k1);
        const double jt1 = L.jt[k1], jd1 = L.jd[k1], jn1 = L.jn[k1];
        if (L.npos(Ib) == 3 && I <= sys_len) {
          const int ib1 = L.inb1_at(L.nposi_dm[Ib] - sys_len - 1, env_len);
          if (!need(ib1, nact_e, "env")) return TJSQ_ERR_ARG;
          if (jt1 != 0.0) {
            add_block_two(P.env.op(ib1, 0), jt1, esub[jc].elec_up, P.merge);
            add_block_two(P.env.op(ib1, 1), jt1, esub[jc].elec_down, P.merge);
          }
          if (jd1 != 0.0) add_block_two(P.env.op(ib1, 2), jd1, esub[jc].spin_sd, P.merge);
          if (jn1 != 0.0) add_block_two(P.env.op(ib1, 3), jn1, esub[jc].num_sn, P.merge);
        }
        if (L.npos(Ib) == 1 && I > sys_len + 1 && C.inb[L.nposi_dm[Ib]] == 0) {
          const int ib1 = L.inb1_at(Ib, sys_len);
          const int jc1 = jc - C.coup_bb[1];
          if (!need(ib1, nact_s, "sys")) return TJSQ_ERR_ARG;
          if (jt1 != 0.0) add_block_two(P.sys.op(ib1, 0), jt1, ssub[jc1].elec_up, P.merge);
          if (jd1 != 0.0) add_block_two(P.sys.op(ib1, 2), jd1, ssub[jc1].spin_sd, P.merge);
          if (jn1 != 0.0) add_block_two(P.sys.op(ib1, 3), jn1, ssub[jc1].num_sn, P.merge);
        }
      }
    }
  }
  P.arena_elems = next;

  // model_wave (get_sys.f90:4127-4240, add_op = 1)
  std::string err;
  const double cone = 1.0;
  LOp s_elec_up = view(so.elec_up), s_num = view(so.num_elec), s_sd = view(so.st_sd), s_si = view(so.st_si);
  em.block_vec_dia(view(P.sys.ham), cone);
  em.vec_block_dia(view(P.env.ham), cone);
  const int is = L.nposi_lat[sys_len + 1], js = L.nposi_lat[sys_len + env_len + 2];
  if (L.esite != 0.0) em.site_vec_site_dia(s_num, s_si, L.esite);
  if (L.esite != 0.0) em.site_vec_site_dia(s_si, s_num, L.esite);
  const int ii = L.neiba[(size_t)is * (L.nsite + 1) + js];
  if (ii != 0) {
    const double jd1 = L.jd[ii], coefx = L.jt[ii], jn1 = L.jn[ii];
    if (coefx != 0.0) em.site_vec_site_ndia(s_elec_up, 'F', s_elec_up, 'F', coefx, err);
    if (jd1 != 0.0) em.site_vec_site_ndia(s_sd, 'B', s_sd, 'B', jd1, err);
    em.site_vec_site_dia(s_num, s_num, jn1);
  }
  if (C.coup_bs[1][1] == 1) {
    em.block_site_vec_ndia(view(ssnc[1].elec_up), 'F', s_elec_up, 'F', cone, err);
    em.block_site_vec_ndia(view(ssnc[1].spin_sd), 'B', s_sd, 'B', cone, err);
    em.block_site_vec_dia(view(ssnc[1].num_sn), s_num, cone);
  }
  if (C.coup_bs[1][2] == 1) {
    em.block_vec_site_ndia(view(ssnc[2].elec_up), 'F', s_elec_up, 'F', cone, err);
    em.block_vec_site_ndia(view(ssnc[2].spin_sd), 'B', s_sd, 'B', cone, err);
    em.block_vec_site_dia(view(ssnc[2].num_sn), s_num, cone);
  }
  if (C.coup_bs[2][1] == 1) {
    em.site_vec_block_ndia(s_elec_up, 'F', trans_view(esnc[1].elec_down), 'F', cone, err);
    em.site_vec_block_ndia(s_sd, 'B', trans_view(esnc[1].spin_sd), 'B', cone, err);
    em.site_vec_block_dia(s_num, view(esnc[1].num_sn), cone);
  }
  if (C.coup_bs[2][2] == 1) {
    em.vec_block_site_ndia(view(esnc[2].elec_up), 'F', s_elec_up, 'F', cone, err);
    em.vec_block_site_ndia(view(esnc[2].spin_sd), 'B', s_sd, 'B', cone, err);
    em.vec_block_site_dia(view(esnc[2].num_sn), s_num, cone);
  }
  {
    int jc = 0;
    for (int i = 1; i <= sys_len + 1 + env_len; ++i) {
      if (C.inb[i] == 0) continue;
      ++jc;
      if (i <= sys_len) {
        const int i1 = L.inb1_at(i, sys_len);
        if (!need(i1, nact_s, "sys")) return TJSQ_ERR_ARG;
        em.block_vec_block_ndia(view(P.sys.op(i1, 0)), 'F', trans_view(esub[jc].elec_down), 'F', cone, err);
        em.block_vec_block_ndia(view(P.sys.op(i1, 2)), 'B', trans_view(esub[jc].spin_sd), 'B', cone, err);
        em.block_vec_block_dia(view(P.sys.op(i1, 3)), view(esub[jc].num_sn), cone);
      } else {
        const int i1 = L.inb1_at(i - sys_len - 1, env_len);
        const int jc1 = jc - C.coup_bb[1];
        if (!need(i1, nact_e, "env")) return TJSQ_ERR_ARG;
        em.block_vec_block_ndia(view(ssub[jc1].elec_up), 'F', trans_view(P.env.op(i1, 1)), 'F', cone, err);
        em.block_vec_block_ndia(view(ssub[jc1].spin_sd), 'B', trans_view(P.env.op(i1, 2)), 'B', cone, err);
        em.block_vec_block_dia(view(ssub[jc1].num_sn), view(P.env.op(i1, 3)), cone);
      }
    }
  }
  if (!err.empty()) { msg = err; return TJSQ_ERR_QN; }
  if (em.bad_dims) { msg = "operator sub-block dimensions disagree with the enlarged bases"; return TJSQ_ERR_ARG; }

  return compile_program(P, shape);
}

// Compiles P.contrib / P.out_blocks into the two-stage task tables.
int compile_program(HostPlan& P, const Shape& shape) {
  auto edge = [&](int M, int N) { return (M > shape.small || N > shape.small) ? shape.big : shape.small; };
  std::vector<Tile> t1big, t1small;
  P.seg1.clear(); P.seg2.clear(); P.tile1.clear(); P.tile2.clear(); P.rtile.clear();
  P.nslots = 0; P.scratch_elems = 0;
  // ---- compile: stage 1 (T = op(A) X of two-sided contributions) ----------
  std::map<std::tuple<int64_t, int, int64_t, int, int>, int64_t> tmap;  // (a_off, a_t, x_off, M, K2) -> T offset
  std::vector<int64_t> t_of(P.contrib.size(), -1);
  std::unordered_map<int64_t, int> agroup;  // stage-1 locality: the A operand
  double fl_ref = 0, fl_exec = 0;
  std::vector<std::pair<int64_t, int64_t>> opref;  // (off, elems) of referenced operator sub-blocks
  // Two-sided contributions into one output block that share the right
  // operand B are factored: sum_j c_j op(A_j) op(X_j) op(B) = T' op(B), with
  // T' formed in stage 1 as one multi-segment product.  lead[q] >= 0 marks
  // a group member; only the group's first member emits a stage-2 segment.
  std::vector<int> lead(P.contrib.size(), -1);
  if (shape.merge_right) {
    std::map<std::tuple<int, int64_t, int, int, int, int>, std::vector<int>> groups;
    for (size_t q = 0; q < P.contrib.size(); ++q) {
      const Contrib& c = P.contrib[q];
      if (c.kind == C_BOTH) groups[std::make_tuple(c.out_block, c.b_off, (int)c.b_t, c.b_ld, c.K2, c.N)].push_back((int)q);
    }
    for (auto& kv : groups) {
      const std::vector<int>& g = kv.second;
      if (g.size() < 2) continue;
      const int64_t toff = P.scratch_elems;
      const Contrib& c0 = P.contrib[g[0]];
      P.scratch_elems += (int64_t)c0.M * c0.K2;
      const int sb = (int)P.seg1.size();
      for (int q : g) {
        const Contrib& c = P.contrib[q];
        lead[q] = g[0];
        Seg s{};
        s.l_off = c.a_off; s.l_ld = c.a_ld; s.l_t = c.a_t; s.l_space = SP_OPS;
        s.r_off = c.x_off; s.r_ld = c.x_ld; s.r_t = c.x_t; s.r_space = SP_X;
        s.K = c.K1; s.c = c.c;
        P.seg1.push_back(s);
        fl_exec += 2.0 * c.M * c.K1 * c.K2;
      }
      const int se = (int)P.seg1.size();
      t_of[g[0]] = toff;
      const int tile = edge(c0.M, c0.K2);
      auto& dst = tile == 64 ? t1big : t1small;
      const int ga = agroup.emplace(c0.a_off, (int)agroup.size()).first->second;
      for (int m0 = 0; m0 < c0.M; m0 += tile)
        for (int n0 = 0; n0 < c0.K2; n0 += tile)
          dst.push_back({toff, c0.M, m0, n0, std::min(tile, c0.M - m0), std::min(tile, c0.K2 - n0), sb, se, SP_T,
                         -1, ga});
    }
  }
  for (size_t q = 0; q < P.contrib.size(); ++q) {
    const Contrib& c = P.contrib[q];
    if (c.kind == C_BOTH && lead[q] >= 0) {
      fl_ref += 2.0 * c.M * c.K1 * c.K2 + 2.0 * c.M * c.K2 * c.N;
      opref.emplace_back(c.a_off, (int64_t)c.M * c.K1);
      opref.emplace_back(c.b_off, (int64_t)c.K2 * c.N);
      continue;
    }
    if (c.kind == C_BOTH) {
      fl_ref += 2.0 * c.M * c.K1 * c.K2 + 2.0 * c.M * c.K2 * c.N;
      auto key = std::make_tuple(c.a_off, (int)c.a_t + 2 * (int)c.x_t, c.x_off, c.M, c.K2);
      auto it = tmap.find(key);
      if (it == tmap.end()) {
        const int64_t toff = P.scratch_elems;
        tmap.emplace(key, toff);
        P.scratch_elems += (int64_t)c.M * c.K2;
        t_of[q] = toff;
        Seg s{};
        s.l_off = c.a_off; s.l_ld = c.a_ld; s.l_t = c.a_t; s.l_space = SP_OPS;
        s.r_off = c.x_off; s.r_ld = c.x_ld; s.r_t = c.x_t; s.r_space = SP_X;
        s.K = c.K1; s.c = 1.0;
        const int sidx = (int)P.seg1.size();
        P.seg1.push_back(s);
        const int tile = edge(c.M, c.K2);
        auto& dst = tile == 64 ? t1big : t1small;   // the kernel's 64x64 items lead the table
        auto ga = agroup.emplace(c.a_off, (int)agroup.size()).first->second;
        for (int m0 = 0; m0 < c.M; m0 += tile)
          for (int n0 = 0; n0 < c.K2; n0 += tile)
            dst.push_back({toff, c.M, m0, n0, std::min(tile, c.M - m0), std::min(tile, c.K2 - n0), sidx, sidx + 1,
                           SP_T, -1, ga});
        fl_exec += 2.0 * c.M * c.K1 * c.K2;
      } else {
        t_of[q] = it->second;
      }
      opref.emplace_back(c.a_off, (int64_t)c.M * c.K1);
      opref.emplace_back(c.b_off, (int64_t)c.K2 * c.N);
    } else if (c.kind == C_LEFT) {
      fl_ref += 2.0 * c.M * c.N * c.K1;
      opref.emplace_back(c.a_off, (int64_t)c.M * c.K1);
    } else if (c.kind == C_RIGHT) {
      fl_ref += 2.0 * c.M * c.N * c.K2;
      opref.emplace_back(c.b_off, (int64_t)c.K2 * c.N);
    } else if (c.kind == C_PROD) {
      fl_ref += 2.0 * c.M * c.N * c.K1;
      if (c.a_space == SP_OPS) opref.emplace_back(c.a_off, (int64_t)c.M * c.K1);
      if (c.b_space == SP_OPS) opref.emplace_back(c.b_off, (int64_t)c.K1 * c.N);
    }
  }


  // ---- stage 2: pull segments per output block ----------------------------
  std::vector<int> order(P.contrib.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = (int)q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return P.contrib[a].out_block < P.contrib[b].out_block; });
  // a factored group emits one segment (its first member's, coefficient 1:
  // the members' coefficients are inside T')
  auto emits = [&](int q) { return lead[q] < 0 || lead[q] == q; };
  std::vector<int> bstart(P.out_blocks.size() + 1, 0);
  for (size_t q = 0; q < P.contrib.size(); ++q)
    if (emits((int)q)) bstart[P.contrib[q].out_block + 1]++;
  for (size_t b = 0; b < P.out_blocks.size(); ++b) bstart[b + 1] += bstart[b];
  {
    std::vector<int> kept;
    kept.reserve(order.size());
    for (int q : order)
      if (emits(q)) kept.push_back(q);
    order.swap(kept);
  }
  P.seg2.resize(order.size());
  for (size_t r = 0; r < order.size(); ++r) {
    const Contrib& c = P.contrib[order[r]];
    Seg s{};
    s.c = lead[order[r]] >= 0 ? 1.0 : c.c;
    switch (c.kind) {
      case C_AXPY:
        s.l_off = c.x_off; s.l_ld = c.x_ld; s.l_space = SP_X; s.r_space = SP_IDENT; s.K = 0;
        break;
      case C_LEFT:
        s.l_off = c.a_off; s.l_ld = c.a_ld; s.l_t = c.a_t; s.l_space = SP_OPS;
        s.r_off = c.x_off; s.r_ld = c.x_ld; s.r_space = SP_X; s.K = c.K1;
        fl_exec += 2.0 * c.M * c.N * c.K1;
        break;
      case C_RIGHT:
        s.l_off = c.x_off; s.l_ld = c.x_ld; s.l_space = SP_X;
        s.r_off = c.b_off; s.r_ld = c.b_ld; s.r_t = c.b_t; s.r_space = SP_OPS; s.K = c.K2;
        fl_exec += 2.0 * c.M * c.N * c.K2;
        break;
      case C_PROD:
        s.l_off = c.a_off; s.l_ld = c.a_ld; s.l_t = c.a_t; s.l_space = c.a_space;
        s.r_off = c.b_off; s.r_ld = c.b_ld; s.r_t = c.b_t; s.r_space = c.b_space; s.K = c.K1;
        fl_exec += 2.0 * c.M * c.N * c.K1;
        break;
      default:
        s.l_off = t_of[order[r]]; s.l_ld = c.M; s.l_space = SP_T;
        s.r_off = c.b_off; s.r_ld = c.b_ld; s.r_t = c.b_t; s.r_space = SP_OPS; s.K = c.K2;
        fl_exec += 2.0 * c.M * c.N * c.K2;
        break;
    }
    P.seg2[r] = s;
  }
  // order 1: each block's GEMM segments sorted by right operand (axpy
  // segments stay first), and blocks visited column strip by column strip
  std::vector<int> border(P.out_blocks.size());
  for (size_t b = 0; b < border.size(); ++b) border[b] = (int)b;
  if (shape.order == 1) {
    for (size_t b = 0; b < P.out_blocks.size(); ++b)
      std::stable_sort(P.seg2.begin() + bstart[b], P.seg2.begin() + bstart[b + 1], [](const Seg& a, const Seg& c) {
        const bool ia = a.r_space == SP_IDENT, ic = c.r_space == SP_IDENT;
        if (ia != ic) return ia;
        if (ia) return false;
        return std::make_tuple((int)a.r_space, a.r_off, (int)a.r_t, a.l_off) <
               std::make_tuple((int)c.r_space, c.r_off, (int)c.r_t, c.l_off);
      });
    std::stable_sort(border.begin(), border.end(),
                     [&](int a, int c) { return P.out_blocks[a].group < P.out_blocks[c].group; });
  }
  // work items: per output block, axpy segments first, then GEMM segments cut
  // into K pieces of at most k_split and grouped into items of bounded work
  std::vector<Seg> pieces;
  std::vector<Tile> t2big, t2small;
  pieces.reserve(P.seg2.size() * 2);
  for (int b : border) {
    const OutBlock& ob = P.out_blocks[b];
    std::vector<std::pair<int, int>> items;  // [begin, end) into pieces
    const int first = (int)pieces.size();
    for (int r = bstart[b]; r < bstart[b + 1]; ++r)
      if (P.seg2[r].r_space == SP_IDENT) pieces.push_back(P.seg2[r]);
    for (int r = bstart[b]; r < bstart[b + 1]; ++r) {
      const Seg& s0 = P.seg2[r];
      if (s0.r_space == SP_IDENT) continue;
      for (int k_lo = 0; k_lo < s0.K; k_lo += shape.k_split) {
        Seg s = s0;
        s.K = std::min(shape.k_split, s0.K - k_lo);
        s.l_off += s0.l_t ? (int64_t)k_lo : (int64_t)k_lo * s0.l_ld;
        s.r_off += s0.r_t ? (int64_t)k_lo * s0.r_ld : (int64_t)k_lo;
        pieces.push_back(s);
      }
    }
    const int last = (int)pieces.size();
    int ib = first, cost = 0;
    for (int q = first; q < last; ++q) {
      const int c = pieces[q].r_space == SP_IDENT ? 1 : (pieces[q].K + shape.kc - 1) / shape.kc;
      if (q > ib && cost + c > shape.item_chunks) {
        items.push_back({ib, q});
        ib = q;
        cost = 0;
      }
      cost += c;
    }
    items.push_back({ib, last});  // also the empty item of a block without terms
    const int tile = edge(ob.M, ob.N);
    const int units = (tile / 32) * (tile / 32);
    auto& dst = tile == 64 ? t2big : t2small;
    for (int m0 = 0; m0 < ob.M; m0 += tile)
      for (int n0 = 0; n0 < ob.N; n0 += tile) {
        const int mt = std::min(tile, ob.M - m0), nt = std::min(tile, ob.N - n0);
        if (items.size() == 1) {
          dst.push_back({ob.off, ob.ld, m0, n0, mt, nt, items[0].first, items[0].second, SP_X, -1, ob.group});
          continue;
        }
        const int s0 = (int)P.nslots;
        for (auto& it : items) {
          dst.push_back({ob.off, ob.ld, m0, n0, mt, nt, it.first, it.second, SP_X, (int)P.nslots, ob.group});
          P.nslots += units;
        }
        // reduce tile: slot = units per partial, out_space = partial leading dimension
        P.rtile.push_back({ob.off, ob.ld, m0, n0, mt, nt, s0, (int)P.nslots, tile, units, ob.group});
      }
  }
  P.seg2.swap(pieces);
  // XCD-aware order: work items of one locality group (operands they share)
  // go to one XCD, whose L2 then serves the repeated operand reads.  Groups
  // are spread over the XCDs largest first onto the least loaded one; within
  // an XCD groups run in that order and items longest first.  Workgroup id g
  // runs on XCD g % xcds, so bin x's p-th item takes id p * xcds + x; bins are
  // padded with empty items (mt = nt = 0, no segments).
  auto xcd_order = [&](std::vector<Tile>& v, const std::vector<Seg>& segs, bool keep) {
    const int NX = std::max(1, shape.xcds);
    std::vector<int64_t> cost(v.size(), 0);
    for (size_t t = 0; t < v.size(); ++t) {
      for (int q = v[t].seg_begin; q < v[t].seg_end; ++q) cost[t] += segs[q].r_space == SP_IDENT ? 1 : segs[q].K;
      cost[t] = std::max<int64_t>(1, cost[t]) * v[t].mt * v[t].nt;
    }
    std::unordered_map<int, int64_t> gcost;
    for (size_t t = 0; t < v.size(); ++t) gcost[v[t].group] += cost[t];
    std::vector<std::pair<int64_t, int>> order;
    for (auto& kv : gcost) order.push_back({kv.second, kv.first});
    std::sort(order.begin(), order.end(), [](const std::pair<int64_t, int>& a, const std::pair<int64_t, int>& b) {
      return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    std::unordered_map<int, int> gbin, grank;
    std::vector<int64_t> load(NX, 0);
    for (size_t r = 0; r < order.size(); ++r) {
      const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      gbin[order[r].second] = b;
      grank[order[r].second] = (int)r;
      load[b] += order[r].first;
    }
    std::vector<int> idx(v.size());
    for (size_t t = 0; t < idx.size(); ++t) idx[t] = (int)t;
    // keep: a group's items stay in table order (the blocks of a column
    // strip walk their operands together); else its longest items run first
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      const int ra = grank[v[a].group], rb = grank[v[b].group];
      return ra != rb ? ra < rb : (!keep && cost[a] > cost[b]);
    });
    std::vector<std::vector<int>> bins(NX);
    for (int t : idx) bins[gbin[v[t].group]].push_back(t);
    size_t maxlen = 0;
    for (auto& b : bins) maxlen = std::max(maxlen, b.size());
    Tile null{0, 1, 0, 0, 0, 0, 0, 0, SP_X, -1, -1};
    std::vector<Tile> out(maxlen * NX, null);
    for (int x = 0; x < NX; ++x)
      for (size_t p = 0; p < bins[x].size(); ++p) out[p * NX + x] = v[bins[x][p]];
    v.swap(out);
  };
  xcd_order(t2big, P.seg2, shape.order == 1);
  xcd_order(t2small, P.seg2, shape.order == 1);
  xcd_order(t1big, P.seg1, false);
  xcd_order(t1small, P.seg1, false);
  P.n64_1 = (int)t1big.size();
  P.tile1 = t1big;
  P.tile1.insert(P.tile1.end(), t1small.begin(), t1small.end());
  P.n64_2 = (int)t2big.size();
  P.tile2 = t2big;
  P.tile2.insert(P.tile2.end(), t2small.begin(), t2small.end());

  std::sort(opref.begin(), opref.end());
  opref.erase(std::unique(opref.begin(), opref.end()), opref.end());
  int64_t op_elems = 0;
  for (auto& p : opref) op_elems += p.second;

  tjsq_plan_stats& st = P.stats;
  st.dim = P.dim;
  st.n_contrib = (int64_t)P.contrib.size();
  st.n_chain = 0;
  for (const Contrib& c : P.contrib) st.n_chain += (c.kind == C_BOTH);
  st.n_stage1 = (int64_t)P.seg1.size();
  st.n_out_blocks = (int64_t)P.out_blocks.size();
  st.n_tiles1 = (int64_t)P.tile1.size();
  st.n_tiles2 = (int64_t)P.tile2.size();
  st.n_slots = P.nslots;
  st.flops_ref = fl_ref;
  st.flops_exec = fl_exec;
  st.op_elems = op_elems;
  st.bytes_fp64 = 8.0 * (2.0 * (double)P.dim + (double)op_elems);
  st.scratch_elems = P.scratch_elems;
  return TJSQ_OK;
}

}  // namespace tjsq
