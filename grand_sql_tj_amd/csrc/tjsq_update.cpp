// tjsq_update.cpp -- density matrix, truncation and block update programs
// (see tjsq_update.hpp for the scheme).
#include "tjsq_update.hpp"

#include <map>
#include <tuple>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace tjsq {

namespace {

constexpr int SU = 2;      // spin step, common.f90:11
constexpr int NEIBT = 12;  // common.f90:34

double par(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

// one placement inside an enlarged sub-block: rows [r0, r0+h), columns
// [c0, c0+w) receive coef * op(source); src < 0 is the identity (h == w)
struct Place {
  int r0, c0, h, w;
  double coef;
  int64_t src;
  int ld;
  uint8_t t;
};
struct SymSub {
  int num_up, num_down, down_dif, row_dim, sdim;
  std::vector<Place> pl;
};
// an enlarged operator (Total_Block) whose sub-blocks are placement sums
struct SymOp {
  bool present = false;
  int len = 0, up_dif = 0, down_dif = 0;
  std::vector<SymSub> sub;
};

// 1x1 site operators of Get_site_operator_su2_U0 (dmrg_sub.f90:370-473) with
// the particle label ub
struct Site {
  int up_dif = 0, down_dif = 0;
  std::vector<OpSub> sub;   // off unused
  std::vector<double> val;
};
struct SiteSet { Site elec_up, elec_down, num_elec, st_sd; };
SiteSet site_ops(int ub) {
  SiteSet s;
  s.elec_up.up_dif = ub; s.elec_up.down_dif = 1;
  s.elec_up.sub = {{0, 0, 1, 1, 1, 0}}; s.elec_up.val = {-std::sqrt(2.0)};
  s.elec_down.up_dif = -ub; s.elec_down.down_dif = 1;
  s.elec_down.sub = {{ub, 1, -1, 1, 1, 0}}; s.elec_down.val = {-std::sqrt(2.0)};
  s.num_elec.sub = {{0, 0, 0, 1, 1, 0}, {ub, 1, 0, 1, 1, 0}}; s.num_elec.val = {0.0, std::sqrt(2.0)};
  s.st_sd.down_dif = SU;
  s.st_sd.sub = {{ub, 1, 0, 1, 1, 0}}; s.st_sd.val = {std::sqrt(1.5)};
  return s;
}

int find_sec(const BasisView& b, int up, int down) {
  for (size_t x = 0; x < b.sec.size(); ++x)
    if (b.sec[x].new_up == up && b.sec[x].new_down == down) return (int)x;
  return -1;
}

struct Ctx {
  std::string err;
  void fail(const std::string& m) { if (err.empty()) err = m; }
};

void place_block(SymSub& d, int r0, int c0, int h, int w, double c, const OpSub& s, bool transposed, Ctx& cx) {
  const int rows = transposed ? s.sdim : s.row_dim, cols = transposed ? s.row_dim : s.sdim;
  if (rows != h || cols != w) { cx.fail("operator sub-block dimensions disagree with the fused basis"); return; }
  d.pl.push_back({r0, c0, h, w, c, s.off, s.row_dim, (uint8_t)transposed});
}

// update_block_dia, block.f90:382-432
SymOp upd_block_dia(const Op& blk, const BasisView& bs, int len, Ctx& cx) {
  SymOp o;
  o.present = true; o.len = len;
  for (const MapSec& k : bs.sec) o.sub.push_back({k.new_up, k.new_down, 0, k.dim, k.dim, {}});
  for (size_t k = 0; k < bs.sec.size(); ++k)
    for (const MapSub& m : bs.sec[k].sub)
      for (const OpSub& s : blk.sub)
        if (s.num_up == m.bl_up && s.num_down == m.bl_down) {
          const double c = std::sqrt((1.0 + bs.sec[k].new_down) / (1.0 + s.num_down));
          place_block(o.sub[k], m.spos, m.spos, m.sdim, m.sdim, c, s, false, cx);
          break;
        }
  return o;
}

// update_site_dia, block.f90:437-498
SymOp upd_site_dia(const Site& st, const BasisView& bs, int len) {
  SymOp o;
  o.present = true; o.len = len;
  for (const MapSec& k : bs.sec) o.sub.push_back({k.new_up, k.new_down, 0, k.dim, k.dim, {}});
  for (size_t k = 0; k < bs.sec.size(); ++k)
    for (const MapSub& m : bs.sec[k].sub)
      for (size_t i = 0; i < st.sub.size(); ++i)
        if (st.sub[i].num_up == m.st_up && st.sub[i].num_down == m.st_down) {
          const double c = std::sqrt((1.0 + bs.sec[k].new_down) / (1.0 + st.sub[i].num_down));
          o.sub[k].pl.push_back({m.spos, m.spos, m.sdim, m.sdim, st.val[i] * c, -1, 0, 0});
          break;
        }
  return o;
}

// sub-block list of the non-diagonal updates (steps <1>-<2> of
// update_block_ndia_*, update_site_ndia_*): ket sector x, bra x + shift
void ndia_structure(const BasisView& bs, int up_dif, int dd1, SymOp& o) {
  for (const MapSec& x : bs.sec)
    for (int dd = -dd1; dd <= dd1; dd += SU) {
      const int y = find_sec(bs, x.new_up + up_dif, x.new_down + dd);
      if (y >= 0) o.sub.push_back({x.new_up, x.new_down, dd, bs.sec[y].dim, x.dim, {}});
    }
}

// update_block_ndia_sys (side 0, block.f90:503-715) and update_block_ndia_env
// (side 1, :720-920; fermion sign from the site's 2S)
SymOp upd_block_ndia(const Op& blk, const BasisView& bs, int len, char flag, int side, Ctx& cx) {
  SymOp o;
  o.present = true; o.len = len; o.up_dif = blk.up_dif; o.down_dif = blk.down_dif;
  const int ud = blk.up_dif, dd1 = blk.down_dif;
  ndia_structure(bs, ud, dd1, o);
  for (SymSub& ns : o.sub) {
    const int rid = find_sec(bs, ns.num_up, ns.num_down);
    const int lid = find_sec(bs, ns.num_up + ud, ns.num_down + ns.down_dif);
    if (rid < 0 || lid < 0) continue;
    const int j1 = ns.num_down + ns.down_dif, j3 = ns.num_down;
    for (const MapSub& x : bs.sec[rid].sub)
      for (int sdd = -dd1; sdd <= dd1; sdd += SU) {
        int bid = -1;
        for (size_t y = 0; y < blk.sub.size(); ++y)
          if (blk.sub[y].num_up == x.bl_up && blk.sub[y].num_down == x.bl_down && blk.sub[y].down_dif == sdd) {
            bid = (int)y;
            break;
          }
        if (bid < 0) continue;
        const int lsu = x.bl_up + ud, lsd = x.bl_down + sdd;
        const MapSub* lm = nullptr;
        for (const MapSub& m : bs.sec[lid].sub)
          if (m.bl_up == lsu && m.bl_down == lsd && m.st_down == x.st_down) { lm = &m; break; }
        if (!lm) continue;
        double c = w6js(j1, dd1, j3, x.bl_down, x.st_down, lsd);
        if (side == 0) {
          c = c * std::sqrt((1.0 + j1) * (1.0 + j3));
          c = c * par((lsd + x.st_down + j3 + dd1) / 2);
        } else {
          c = c * par((lsd + x.st_down + j3 + dd1) / 2) * std::sqrt((1.0 + j1) * (1.0 + j3));
          if (c != 0.0) c = c * (flag == 'F' && x.st_down % 2 != 0 ? -1.0 : 1.0);
        }
        if (c != 0.0) place_block(ns, lm->spos, x.spos, lm->sdim, x.sdim, c, blk.sub[bid], false, cx);
      }
  }
  return o;
}

// update_site_ndia_sys (side 0, block.f90:925-1112; fermion sign from the
// block's 2S) and update_site_ndia_env (side 1, :1113-1305)
SymOp upd_site_ndia(const Site& st, const BasisView& bs, int len, char flag, int side) {
  SymOp o;
  o.present = true; o.len = len; o.up_dif = st.up_dif; o.down_dif = st.down_dif;
  const int ud = st.up_dif, dd1 = st.down_dif;
  ndia_structure(bs, ud, dd1, o);
  for (SymSub& ns : o.sub) {
    const int rid = find_sec(bs, ns.num_up, ns.num_down);
    const int lid = find_sec(bs, ns.num_up + ud, ns.num_down + ns.down_dif);
    if (rid < 0 || lid < 0) continue;
    const int j1 = ns.num_down + ns.down_dif, j3 = ns.num_down;
    for (const MapSub& x : bs.sec[rid].sub)
      for (int sdd = -dd1; sdd <= dd1; sdd += SU) {
        int sid = -1;
        for (size_t y = 0; y < st.sub.size(); ++y)
          if (st.sub[y].num_up == x.st_up && st.sub[y].num_down == x.st_down && st.sub[y].down_dif == sdd) {
            sid = (int)y;
            break;
          }
        if (sid < 0) continue;
        const int lsu = x.st_up + ud, lsd = x.st_down + sdd;
        const MapSub* lm = nullptr;
        for (const MapSub& m : bs.sec[lid].sub)
          if (m.st_up == lsu && m.st_down == lsd && m.bl_down == x.bl_down) { lm = &m; break; }
        if (!lm) continue;
        double c = w6js(j1, dd1, j3, x.st_down, x.bl_down, lsd);
        if (side == 0) {
          c = c * par((x.bl_down + x.st_down + j1 + dd1) / 2) * std::sqrt((1.0 + j1) * (1.0 + j3));
          if (c != 0.0) c = c * (flag == 'F' && x.bl_down % 2 != 0 ? -1.0 : 1.0);
        } else {
          c = c * std::sqrt((1.0 + j1) * (1.0 + j3)) * par((x.bl_down + x.st_down + j1 + dd1) / 2);
        }
        if (c != 0.0) ns.pl.push_back({lm->spos, x.spos, x.sdim, x.sdim, c * st.val[sid], -1, 0, 0});
      }
  }
  return o;
}

// Get_hamiltonian_ndia, get_sys.f90:583-760: block (x) site bond term into H,
// with the h.c. written explicitly for fermions (site down_dif = 1)
void ham_ndia(const Op& blk, const Site& st, const BasisView& bs, double coef, SymOp& H, char fsign, int side,
              Ctx& cx) {
  const int ud = blk.up_dif, dd1 = blk.down_dif;
  for (SymSub& hs : H.sub) {
    const int rid = find_sec(bs, hs.num_up, hs.num_down);
    const int lid = find_sec(bs, hs.num_up, hs.num_down + hs.down_dif);
    if (rid < 0 || lid < 0) continue;
    for (const MapSub& x : bs.sec[rid].sub)
      for (int sdd = -dd1; sdd <= dd1; sdd += SU) {
        int bid = -1;
        for (size_t y = 0; y < blk.sub.size(); ++y)
          if (blk.sub[y].num_up == x.bl_up && blk.sub[y].num_down == x.bl_down && blk.sub[y].down_dif == sdd) {
            bid = (int)y;
            break;
          }
        if (bid < 0) continue;
        const int lsu = x.bl_up + ud, lsd = x.bl_down + sdd;
        for (int nsd = std::abs(x.st_down - st.down_dif); nsd <= x.st_down + st.down_dif; nsd += SU) {
          const int st_dd = x.st_down - nsd, nsu = x.st_up - st.up_dif;
          int sid = -1;
          for (size_t y = 0; y < st.sub.size(); ++y)
            if (st.sub[y].num_up == nsu && st.sub[y].num_down == nsd && st.sub[y].down_dif == st_dd) {
              sid = (int)y;
              break;
            }
          const MapSub* lm = nullptr;
          for (const MapSub& m : bs.sec[lid].sub)
            if (m.bl_up == lsu && m.bl_down == lsd && m.st_down == nsd) { lm = &m; break; }
          if (!lm || sid < 0) continue;
          const int j1 = lsd, j2 = x.bl_down, j3 = dd1, j4 = x.st_down, j5 = nsd, j6 = hs.num_down;
          double c1 = w6js(j1, j2, j3, j4, j5, j6);
          c1 = c1 * std::sqrt((1.0 + j6) / (1.0 + j3)) * par((j2 + j5 + j3 + j6) / 2) * st.val[sid];
          double sign = 1.0;
          if (side == 0 && fsign == 'F') {
            if (x.bl_up % 2 != 0) sign = -1.0;
            if (x.bl_down % 2 != 0) sign = -1.0;
          }
          if (c1 == 0.0) continue;
          c1 = c1 * coef * sign;
          place_block(hs, lm->spos, x.spos, lm->sdim, x.sdim, c1, blk.sub[bid], false, cx);
          if (st.down_dif == 1) place_block(hs, x.spos, lm->spos, x.sdim, lm->sdim, c1, blk.sub[bid], true, cx);
        }
      }
  }
}

// Get_hamiltonian_dia, get_sys.f90:506-577 (FlagConjg false, 'N','N'):
// H += coef / sqrt(1 + 2S) * A * B per sector.  A and B are embeddings over
// the same fused-basis grid; the product is formed symbolically when one
// factor of every overlapping pair is an identity placement.
void ham_dia(const SymOp& A, const SymOp& B, double coef, SymOp& H, Ctx& cx) {
  for (SymSub& hs : H.sub) {
    int a = -1, b = -1;
    for (size_t x = 0; x < A.sub.size(); ++x)
      if (A.sub[x].num_up == hs.num_up && A.sub[x].num_down == hs.num_down) { a = (int)x; break; }
    for (size_t x = 0; x < B.sub.size(); ++x)
      if (B.sub[x].num_up == hs.num_up && B.sub[x].num_down == hs.num_down) { b = (int)x; break; }
    if (a < 0 || b < 0) continue;
    const double c1 = coef / std::sqrt(1.0 + hs.num_down);
    for (const Place& p : A.sub[a].pl)
      for (const Place& q : B.sub[b].pl) {
        if (p.c0 + p.w <= q.r0 || q.r0 + q.h <= p.c0) continue;
        if (p.c0 != q.r0 || p.w != q.h) { cx.fail("operator product over misaligned placements"); return; }
        if (p.src >= 0 && q.src >= 0) { cx.fail("operator product of two data blocks is not supported"); return; }
        const double c = c1 * p.coef * q.coef;
        if (c == 0.0) continue;
        if (q.src < 0)
          hs.pl.push_back({p.r0, q.c0, p.h, q.w, c, p.src, p.ld, p.t});
        else
          hs.pl.push_back({p.r0, q.c0, p.h, q.w, c, q.src, q.ld, q.t});
      }
  }
}

// block_add_block_two, block.f90:1782-1814 (forward-only sub-block search)
void add_two(const SymOp& b1, double coef, SymOp& nb) {
  size_t x1 = 0;
  for (SymSub& d : nb.sub)
    for (size_t x = x1; x < b1.sub.size(); ++x) {
      const SymSub& s = b1.sub[x];
      if (s.num_up == d.num_up && s.num_down == d.num_down && s.down_dif == d.down_dif) {
        for (Place p : s.pl) {
          p.coef *= coef;
          d.pl.push_back(p);
        }
        x1 = x + 1;
        break;
      }
    }
}

// a stored operator as a SymOp (every sub-block one placement)
SymOp as_sym(const Op& o) {
  SymOp s;
  s.present = true; s.len = o.len; s.up_dif = o.up_dif; s.down_dif = o.down_dif;
  for (const OpSub& q : o.sub)
    s.sub.push_back({q.num_up, q.num_down, q.down_dif, q.row_dim, q.sdim,
                     {{0, 0, q.row_dim, q.sdim, 1.0, q.off, q.row_dim, 0}}});
  return s;
}

int read_op(const tjsq_opset& in, int op, int first, Op& o, std::string& msg) {
  const int* h = in.hdr + 4 * op;
  o.len = h[0]; o.up_dif = h[1]; o.down_dif = h[2];
  o.sub.resize(h[3]);
  for (int s = 0; s < h[3]; ++s) {
    const int* q = in.sub + 5 * (first + s);
    OpSub& d = o.sub[s];
    d.num_up = q[0]; d.num_down = q[1]; d.down_dif = q[2]; d.row_dim = q[3]; d.sdim = q[4];
    d.off = in.off[first + s];
    if (d.row_dim < 0 || d.sdim < 0 || d.off < 0 || d.off + (int64_t)d.row_dim * d.sdim > in.ndata) {
      msg = "operator sub-block outside the data array";
      return TJSQ_ERR_ARG;
    }
  }
  return TJSQ_OK;
}

int read_ops(const tjsq_opset& in, std::vector<Op>& ops, std::string& msg) {
  if (in.nops < 1 || !in.hdr || !in.sub || !in.off) { msg = "empty operator set"; return TJSQ_ERR_ARG; }
  ops.resize(in.nops);
  int first = 0;
  for (int op = 0; op < in.nops; ++op) {
    int rc = read_op(in, op, first, ops[op], msg);
    if (rc) return rc;
    first += in.hdr[4 * op + 3];
  }
  return TJSQ_OK;
}

void read_basis(const tjsq_basis& b, BasisView& v) {
  v.sec.resize(b.nsec);
  const int* q = b.sub;
  for (int k = 0; k < b.nsec; ++k) {
    MapSec& s = v.sec[k];
    s.new_up = b.sec[4 * k]; s.new_down = b.sec[4 * k + 1]; s.dim = b.sec[4 * k + 2];
    s.sub.resize(b.sec[4 * k + 3]);
    for (auto& m : s.sub) { m = {q[0], q[1], q[2], q[3], q[4], q[5]}; q += 6; }
  }
}

int64_t max_extent(const tjsq_opset& in) {
  int64_t e = 0;
  int nsub = 0;
  for (int op = 0; op < in.nops; ++op) nsub += in.hdr[4 * op + 3];
  for (int s = 0; s < nsub; ++s) e = std::max(e, in.off[s] + (int64_t)in.sub[5 * s + 3] * in.sub[5 * s + 4]);
  return e;
}

// Projection of one (symbolic) operator with the truncation operator T:
// update_trun_dia (block.f90:1311-1358) or update_trun_ndia (:1360-1503).
// Appends the output operator to pg.out and its contributions to pg.hp.
struct Projector {
  Program& pg;
  const Op& T;
  Ctx& cx;
  int find_t(int up, int down) const {
    for (size_t x = 0; x < T.sub.size(); ++x)
      if (T.sub[x].num_up == up && T.sub[x].num_down == down) return (int)x;
    return -1;
  }
  int new_block(int row, int col, int group) {
    pg.hp.out_blocks.push_back({pg.out.ndata, std::max(1, row), row, col, group});
    pg.out.off.push_back(pg.out.ndata);
    pg.out.ndata += (int64_t)row * col;
    return (int)pg.hp.out_blocks.size() - 1;
  }
  // every placement of ori sub-block s projected into output block ob
  void pieces(const SymSub& s, const OpSub& tl, const OpSub& tr, int ob) {
    if (s.row_dim != tl.row_dim || s.sdim != tr.row_dim) {
      cx.fail("truncation operator rows disagree with the operator sub-block");
      return;
    }
    for (const Place& p : s.pl) {
      if (p.coef == 0.0 || tl.sdim == 0 || tr.sdim == 0 || p.h == 0 || p.w == 0) continue;
      Contrib q{};
      q.out_block = ob;
      q.c = p.coef;
      q.M = tl.sdim;
      q.N = tr.sdim;
      q.a_off = tl.off + p.r0; q.a_ld = tl.row_dim; q.a_t = 1;   // T_lhs[rows]^T
      q.b_off = tr.off + p.c0; q.b_ld = tr.row_dim; q.b_t = 0;   // T_rhs[cols]
      if (p.src < 0) {
        q.kind = C_PROD;
        q.a_space = SP_OPS; q.b_space = SP_OPS;
        q.K1 = p.h;
      } else {
        q.kind = C_BOTH;
        q.x_off = p.src; q.x_ld = p.ld; q.x_t = p.t;
        q.K1 = p.h; q.K2 = p.w;
      }
      pg.hp.contrib.push_back(q);
    }
  }
  void project(const SymOp& ori, bool dia) {
    const int first_sub = (int)pg.out.sub.size() / 5;
    int nsub = 0;
    const int gbase = (int)pg.hp.out_blocks.size();
    if (dia) {
      for (size_t i = 0; i < T.sub.size(); ++i) {
        const OpSub& t = T.sub[i];
        pg.out.sub.insert(pg.out.sub.end(), {t.num_up, t.num_down, 0, t.sdim, t.sdim});
        const int ob = new_block(t.sdim, t.sdim, gbase + (int)i);
        ++nsub;
        for (const SymSub& s : ori.sub)
          if (s.num_up == t.num_up && s.num_down == t.num_down) {
            pieces(s, t, t, ob);
            break;
          }
      }
    } else {
      const int ud = ori.up_dif, dd1 = ori.down_dif;
      for (size_t x = 0; x < T.sub.size(); ++x)
        for (int dd = -dd1; dd <= dd1; dd += SU) {
          const OpSub& tr = T.sub[x];
          const int ly = find_t(tr.num_up + ud, tr.num_down + dd);
          if (ly < 0) continue;
          const OpSub& tl = T.sub[ly];
          pg.out.sub.insert(pg.out.sub.end(), {tr.num_up, tr.num_down, dd, tl.sdim, tr.sdim});
          const int ob = new_block(tl.sdim, tr.sdim, gbase + (int)x);
          ++nsub;
          const int rid = find_t(tr.num_up, tr.num_down);
          if (rid < 0) continue;
          for (const SymSub& s : ori.sub)
            if (s.num_up == tr.num_up && s.num_down == tr.num_down && s.down_dif == dd) {
              pieces(s, tl, T.sub[rid], ob);
              break;
            }
        }
    }
    (void)first_sub;
    pg.out.hdr.insert(pg.out.hdr.end(), {ori.len, ori.up_dif, ori.down_dif, nsub});
    pg.out.nops++;
  }
};

int finish(Program& pg, const Shape& shape, Ctx& cx, std::string& msg) {
  if (!cx.err.empty()) { msg = cx.err; return TJSQ_ERR_QN; }
  int rc = compile_program(pg.hp, shape);
  if (rc) { msg = "program compilation failed"; return rc; }
  pg.hp.dim = pg.out.ndata;
  pg.hp.stats.dim = pg.out.ndata;
  return TJSQ_OK;
}

}  // namespace

// --------------------------------------------------------------------------
// Get_Hamiltonian_Sys (get_sys.f90:4-246) / Get_Hamiltonian_Env (:247-478)
// --------------------------------------------------------------------------
int build_update_program(const tjsq_lattice& lat, int side, int sys_len, int env_len, const tjsq_opset& pre,
                         const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                         std::string& msg) {
  pg = Program();
  if (side != 0 && side != 1) { msg = "side must be 0 (system) or 1 (environment)"; return TJSQ_ERR_ARG; }
  if (lat.ny < 1 || lat.nx_tables < 1 || sys_len < 1 || env_len < 1 ||
      sys_len + env_len + 2 > lat.ny * lat.nx_tables) {
    msg = "lattice too small for sys_len + env_len + 2 sites";
    return TJSQ_ERR_ARG;
  }
  std::vector<Op> ops, tops;
  int rc = read_ops(pre, ops, msg);
  if (rc) return rc;
  if ((pre.nops - 1) % 4 != 0) { msg = "operator count must be 1 + 4 * active sites"; return TJSQ_ERR_ARG; }
  if ((rc = read_ops(trun, tops, msg))) return rc;
  if (trun.nops != 1) { msg = "the truncation operator set must hold one operator"; return TJSQ_ERR_ARG; }
  const Op& T = tops[0];
  BasisView bs;
  read_basis(basis, bs);

  LatticeTables L;
  build_lattice(lat, L);
  lattice_setup(L, lat.ny, sys_len, env_len);
  const int pre_len = side == 0 ? sys_len : env_len;
  if (pre.hdr[0] != pre_len) { msg = "block length disagrees with sys_len / env_len"; return TJSQ_ERR_ARG; }
  if (pre_len + 1 > L.nsite - 2) { msg = "block too long for the lattice tables"; return TJSQ_ERR_ARG; }
  const int nact = (pre.nops - 1) / 4;
  const int nleg_pre = L.nleg11[pre_len], nleg1 = L.nleg11[pre_len + 1];
  if (nact < nleg_pre) { msg = "block model has fewer active sites than nleg11(len)"; return TJSQ_ERR_ARG; }
  const int mshift = sys_len + 1;  // environment chain offset (dmrg_sub.f90:692, 989)
  const int new_len = pre_len + 1;
  const int jnew = L.inb1_at(pre_len + 1, pre_len + 1);
  const int nslot = std::max({nleg_pre, nleg1, jnew < 10000 ? jnew : 0}) + 1;
  auto op_of = [&](int i1, int q) -> const Op& { return ops[1 + 4 * (i1 - 1) + q]; };
  const SiteSet st = site_ops(lat.up_bias);
  Ctx cx;

  // neighbours of the new site inside the block: (active index i1, bond k)
  std::vector<std::pair<int, int>> nb;
  for (int k = 1; k <= NEIBT; ++k) {
    int II;
    if (side == 0) {
      II = L.neib_at(pre_len + 1, k);
    } else {
      const int I = L.neib_at(L.nposi_lat[pre_len + 1 + mshift], k);
      II = L.nposi_dm[I] - mshift;
    }
    if (II > pre_len || II <= 0) continue;
    const int i1 = L.inb1_at(II, pre_len);
    if (i1 > 0 && i1 <= nleg_pre) nb.push_back({i1, k});
  }

  SymOp H = upd_block_dia(ops[0], bs, new_len, cx);
  // next%sub(i)%{elec_up, elec_down, spin_sd, num_sn}
  std::vector<SymOp> nxt[4];
  for (auto& v : nxt) v.assign(nslot, SymOp());
  for (int ij = 1; ij <= 5; ij += 2) {
    for (auto& [i1, k] : nb) {
      if (ij == 1 && L.jt[k] != 0.0) ham_ndia(op_of(i1, 0), st.elec_up, bs, L.jt[k], H, 'F', side, cx);
      if (ij == 3 && L.jd[k] != 0.0) ham_ndia(op_of(i1, 2), st.st_sd, bs, L.jd[k], H, 'B', side, cx);
    }
    for (int y = 1; y <= pre_len; ++y) {
      const int I1 = L.inb1_at(y, pre_len);
      if (I1 > nleg_pre) continue;
      if (ij == 1) {
        nxt[0][I1] = upd_block_ndia(op_of(I1, 0), bs, new_len, 'F', side, cx);
        nxt[1][I1] = upd_block_ndia(op_of(I1, 1), bs, new_len, 'F', side, cx);
      } else if (ij == 3) {
        nxt[2][I1] = upd_block_ndia(op_of(I1, 2), bs, new_len, 'B', side, cx);
      } else {
        nxt[3][I1] = upd_block_dia(op_of(I1, 3), bs, new_len, cx);
      }
    }
    SymOp tmp, tmp1;
    if (ij == 1) {
      tmp = upd_site_ndia(st.elec_up, bs, new_len, 'F', side);
      tmp1 = upd_site_ndia(st.elec_down, bs, new_len, 'F', side);
    } else if (ij == 3) {
      tmp = upd_site_ndia(st.st_sd, bs, new_len, 'B', side);
    } else {
      tmp = upd_site_dia(st.num_elec, bs, new_len);
      for (auto& [i1, k] : nb) ham_dia(nxt[3][i1], tmp, L.jn[k], H, cx);
      if (L.esite != 0.0) add_two(tmp, L.esite, H);
    }
    // move the embedded operators to their indices at length L + 1
    const int q0 = ij == 1 ? 0 : (ij == 3 ? 2 : 3);
    const int qn = ij == 1 ? 2 : 1;
    for (int I = 1; I <= pre_len; ++I) {
      const int I1 = L.inb1_at(I, pre_len), J = L.inb1_at(I, pre_len + 1);
      if (J <= nleg1 && J != I1) {
        if (I1 >= nslot) { msg = "active-site table refers past the block's sites"; return TJSQ_ERR_ARG; }
        for (int q = q0; q < q0 + qn; ++q) nxt[q][J] = nxt[q][I1];
      }
    }
    if (jnew <= 0 || jnew >= nslot) { msg = "new site is not active"; return TJSQ_ERR_ARG; }
    nxt[q0][jnew] = tmp;
    if (ij == 1) nxt[1][jnew] = tmp1;
    for (int j = nleg1 + 1; j <= nleg_pre; ++j)
      for (int q = q0; q < q0 + qn; ++q) nxt[q][j] = SymOp();
  }

  // projection with T (update_trun_dia / update_trun_ndia), output model order
  Projector pj{pg, T, cx};
  pj.project(H, true);
  for (int ii = 1; ii <= nleg1; ++ii)
    for (int q = 0; q < 4; ++q) {
      if (!nxt[q][ii].present) { msg = "active-site operator missing after the enlargement"; return TJSQ_ERR_ARG; }
      pj.project(nxt[q][ii], q == 3);
    }
  pg.x_elems = max_extent(pre);
  pg.ops_elems = max_extent(trun);
  return finish(pg, shape, cx, msg);
}

int build_projection_program(const tjsq_opset& ops_in, const int* ndia, const tjsq_opset& trun, const Shape& shape,
                             Program& pg, std::string& msg) {
  pg = Program();
  std::vector<Op> ops, tops;
  int rc = read_ops(ops_in, ops, msg);
  if (rc) return rc;
  if ((rc = read_ops(trun, tops, msg))) return rc;
  if (trun.nops != 1) { msg = "the truncation operator set must hold one operator"; return TJSQ_ERR_ARG; }
  Ctx cx;
  Projector pj{pg, tops[0], cx};
  for (size_t i = 0; i < ops.size(); ++i) pj.project(as_sym(ops[i]), !(ndia && ndia[i]));
  pg.x_elems = max_extent(ops_in);
  pg.ops_elems = max_extent(trun);
  return finish(pg, shape, cx, msg);
}

namespace {

// a site operator given as a one-operator set of 1x1 sub-blocks with host
// data (the layout of Get_site_operator's Total_Block, dmrg_sub.f90:370-473)
int read_site(const tjsq_opset& in, Site& st, std::string& msg) {
  std::vector<Op> ops;
  int rc = read_ops(in, ops, msg);
  if (rc) return rc;
  if (in.nops != 1 || !in.data || in.data_on_device) {
    msg = "the site operator must be one operator with host data";
    return TJSQ_ERR_ARG;
  }
  st = Site();
  st.up_dif = ops[0].up_dif;
  st.down_dif = ops[0].down_dif;
  for (const OpSub& q : ops[0].sub) {
    if (q.row_dim != 1 || q.sdim != 1) { msg = "site operator sub-blocks must be 1 x 1"; return TJSQ_ERR_ARG; }
    st.sub.push_back(q);
    st.val.push_back(in.data[q.off]);
  }
  return TJSQ_OK;
}

bool all_dia(const std::vector<Op>& ops) {
  for (const Op& o : ops) {
    if (o.up_dif != 0 || o.down_dif != 0) return false;
    for (const OpSub& q : o.sub)
      if (q.down_dif != 0 || q.row_dim != q.sdim) return false;
  }
  return true;
}

}  // namespace

// One step of Get_sys_oper_dia / Get_env_oper_dia (get_sys.f90:764-891):
// every diagonal operator of the block at length L embedded into the fused
// basis (update_block_dia), the new site's operator embedded
// (update_site_dia), each projected with T (update_trun_dia).
namespace {

// the product A (x) s of a diagonal block operator A and the new site's
// operator s in the fused basis, as a reduced operator: per fused sector k
// and sub-sector x (block part bl, site part st) the placement
// sqrt((1 + 2S_k) / (1 + 2S_bl)) * s(st) / sqrt(1 + 2S_st) * A(bl), the
// embedding of update_block_dia times the site operator's value
SymOp upd_block_site_product(const Op& blk, const Site& st, const BasisView& bs, int len, Ctx& cx) {
  SymOp o;
  o.present = true; o.len = len;
  for (const MapSec& k : bs.sec) o.sub.push_back({k.new_up, k.new_down, 0, k.dim, k.dim, {}});
  for (size_t k = 0; k < bs.sec.size(); ++k)
    for (const MapSub& m : bs.sec[k].sub) {
      double sv = 0.0;
      for (size_t i = 0; i < st.sub.size(); ++i)
        if (st.sub[i].num_up == m.st_up && st.sub[i].num_down == m.st_down) {
          sv = st.val[i] / std::sqrt(1.0 + st.sub[i].num_down);
          break;
        }
      if (sv == 0.0) continue;
      for (const OpSub& q : blk.sub)
        if (q.num_up == m.bl_up && q.num_down == m.bl_down) {
          const double c = std::sqrt((1.0 + bs.sec[k].new_down) / (1.0 + q.num_down)) * sv;
          place_block(o.sub[k], m.spos, m.spos, m.sdim, m.sdim, c, q, false, cx);
          break;
        }
    }
  return o;
}

}  // namespace

int build_dia_transport_program(const tjsq_opset* ops_in, const tjsq_opset* site, const tjsq_basis& basis,
                                const tjsq_opset& trun, const Shape& shape, Program& pg, std::string& msg) {
  return build_dia_transport_program(ops_in, site, 0, basis, trun, shape, pg, msg);
}

int build_dia_transport_program(const tjsq_opset* ops_in, const tjsq_opset* site, int nprod, const tjsq_basis& basis,
                                const tjsq_opset& trun, const Shape& shape, Program& pg, std::string& msg) {
  pg = Program();
  std::vector<Op> ops, tops;
  int rc;
  if (ops_in && ops_in->nops > 0) {
    if ((rc = read_ops(*ops_in, ops, msg))) return rc;
    if (!all_dia(ops)) { msg = "transported operators must be diagonal (square, down_dif 0)"; return TJSQ_ERR_ARG; }
  }
  Site st;
  if (site && (rc = read_site(*site, st, msg))) return rc;
  if (site && (st.up_dif != 0 || st.down_dif != 0)) { msg = "the site operator must be diagonal"; return TJSQ_ERR_ARG; }
  if (ops.empty() && !site) { msg = "nothing to transport"; return TJSQ_ERR_ARG; }
  if ((rc = read_ops(trun, tops, msg))) return rc;
  if (trun.nops != 1) { msg = "the truncation operator set must hold one operator"; return TJSQ_ERR_ARG; }
  BasisView bs;
  read_basis(basis, bs);
  Ctx cx;
  Projector pj{pg, tops[0], cx};
  if (nprod < 0 || nprod > (int)ops.size() || (nprod > 0 && !site)) {
    msg = "product count must lie within the operators and needs the site operator";
    return TJSQ_ERR_ARG;
  }
  // output order: ops[0, nprod), the site operator, ops[nprod, end), then
  // the products (with nprod = 0: ops, then the site operator)
  for (int i = 0; i < nprod; ++i) pj.project(upd_block_dia(ops[i], bs, basis.length, cx), true);
  if (site && nprod > 0) pj.project(upd_site_dia(st, bs, basis.length), true);
  for (size_t i = nprod; i < ops.size(); ++i) pj.project(upd_block_dia(ops[i], bs, basis.length, cx), true);
  if (site && nprod == 0) pj.project(upd_site_dia(st, bs, basis.length), true);
  for (int i = 0; i < nprod; ++i) pj.project(upd_block_site_product(ops[i], st, bs, basis.length, cx), true);
  pg.x_elems = ops.empty() ? 0 : max_extent(*ops_in);
  pg.ops_elems = max_extent(trun);
  return finish(pg, shape, cx, msg);
}

// measure_sys_block / measure_sys_site (side 0, measure.f90:1241-1355) and
// measure_env_block / measure_env_site (side 1, :1359-1470) as products
// laid out like the wavefunction: output copy j (of nops + 1, each D long)
// holds Y_j with <psi|O_j|psi> = psi . Y_j, where per wavefunction sector
// and fused sub-sector x with block part bl (site part st)
//   side 0: Y_j[rows of x]    = O_j[bl] psi[rows of x] / sqrt(1 + 2S_bl)
//   side 1: Y_j[columns of x] = psi[columns of x] O_j[bl] / sqrt(1 + 2S_bl)
// and the last copy Y_site = site[st] psi[x] / sqrt(1 + 2S_st).  Entries no
// term reaches are not written (the caller clears y).
int build_measure_program(int side, const tjsq_opset* ops_in, const tjsq_opset* site, const tjsq_basis& basis,
                          int nw, const int* wsec, const Shape& shape, Program& pg, std::string& msg) {
  pg = Program();
  if ((side != 0 && side != 1) || nw < 1 || !wsec) { msg = "bad measurement arguments"; return TJSQ_ERR_ARG; }
  std::vector<Op> ops;
  int rc;
  if (ops_in && ops_in->nops > 0) {
    if ((rc = read_ops(*ops_in, ops, msg))) return rc;
    if (!all_dia(ops)) { msg = "measured block operators must be diagonal (square, down_dif 0)"; return TJSQ_ERR_ARG; }
  }
  Site st;
  if (site && (rc = read_site(*site, st, msg))) return rc;
  BasisView bs;
  read_basis(basis, bs);
  int64_t D = 0;
  std::vector<int64_t> woff(nw);
  for (int i = 0; i < nw; ++i) {
    const int* w = wsec + 6 * i;
    if (w[4] < 0 || w[5] < 0) { msg = "negative sector dimension"; return TJSQ_ERR_ARG; }
    woff[i] = D;
    D += (int64_t)w[4] * w[5];
  }
  const int ncopy = (int)ops.size() + 1;
  auto out_block = [&](int64_t off, int ld, int M, int N) {
    pg.hp.out_blocks.push_back({off, std::max(1, ld), M, N, (int)pg.hp.out_blocks.size()});
    return (int)pg.hp.out_blocks.size() - 1;
  };
  for (int i = 0; i < nw; ++i) {
    const int* w = wsec + 6 * i;
    const int ds = w[4], de = w[5];
    if (ds == 0 || de == 0) continue;
    const int k = find_sec(bs, side == 0 ? w[0] : w[2], side == 0 ? w[1] : w[3]);
    if (k < 0) continue;
    if (bs.sec[k].dim != (side == 0 ? ds : de)) { msg = "fused basis sector disagrees with the wavefunction"; return TJSQ_ERR_ARG; }
    for (const MapSub& m : bs.sec[k].sub) {
      if (m.sdim == 0) continue;
      // the rows (side 0) or columns (side 1) of sub-sector x inside sector i
      const int64_t xo = woff[i] + (side == 0 ? (int64_t)m.spos : (int64_t)m.spos * ds);
      const int M = side == 0 ? m.sdim : ds, N = side == 0 ? de : m.sdim;
      for (size_t j = 0; j < ops.size(); ++j) {
        for (const OpSub& q : ops[j].sub) {
          if (q.num_up != m.bl_up || q.num_down != m.bl_down) continue;
          if (q.row_dim != m.sdim) { msg = "operator sub-block disagrees with the fused basis"; return TJSQ_ERR_ARG; }
          Contrib c{};
          c.kind = side == 0 ? C_LEFT : C_RIGHT;
          c.out_block = out_block((int64_t)j * D + xo, ds, M, N);
          c.x_off = xo; c.x_ld = ds;
          c.M = M; c.N = N;
          c.c = 1.0 / std::sqrt(1.0 + m.bl_down);
          if (side == 0) { c.a_off = q.off; c.a_ld = q.row_dim; c.a_t = 0; c.K1 = m.sdim; }
          else { c.b_off = q.off; c.b_ld = q.row_dim; c.b_t = 0; c.K2 = m.sdim; }
          pg.hp.contrib.push_back(c);
          break;
        }
      }
      if (site) {
        for (size_t y = 0; y < st.sub.size(); ++y) {
          if (st.sub[y].num_up != m.st_up || st.sub[y].num_down != m.st_down) continue;
          Contrib c{};
          c.kind = C_AXPY;
          c.out_block = out_block((int64_t)(ncopy - 1) * D + xo, ds, M, N);
          c.x_off = xo; c.x_ld = ds;
          c.M = M; c.N = N;
          c.c = st.val[y] / std::sqrt(1.0 + m.st_down);
          if (c.c != 0.0) pg.hp.contrib.push_back(c);
          else pg.hp.out_blocks.pop_back();
          break;
        }
      }
    }
  }
  const int len = basis.length;
  pg.out.nops = ncopy;
  for (int j = 0; j < ncopy; ++j) pg.out.hdr.insert(pg.out.hdr.end(), {len, 0, 0, 0});
  pg.out.ndata = (int64_t)ncopy * D;
  pg.x_elems = D;
  pg.ops_elems = ops.empty() ? 0 : max_extent(*ops_in);
  Ctx cx;
  return finish(pg, shape, cx, msg);
}

// One step of Get_sys_oper_ndia / Get_env_oper_ndia and Change_*_oper_ndia
// (get_sys.f90:998-1238): every operator of the block at length L embedded
// into the fused basis (update_block_ndia_sys / _env), the new site's
// operator embedded (update_site_ndia_sys / _env), each projected with T
// (update_trun_ndia); flag 'B' or 'F' selects the fermion signs of the
// embeddings.
int build_ndia_transport_program(int side, char flag, const tjsq_opset* ops_in, const tjsq_opset* site,
                                 const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                                 std::string& msg) {
  pg = Program();
  if ((side != 0 && side != 1) || (flag != 'B' && flag != 'F')) { msg = "bad side or sign flag"; return TJSQ_ERR_ARG; }
  std::vector<Op> ops, tops;
  int rc;
  if (ops_in && ops_in->nops > 0 && (rc = read_ops(*ops_in, ops, msg))) return rc;
  Site st;
  if (site && (rc = read_site(*site, st, msg))) return rc;
  if (ops.empty() && !site) { msg = "nothing to transport"; return TJSQ_ERR_ARG; }
  if ((rc = read_ops(trun, tops, msg))) return rc;
  if (trun.nops != 1) { msg = "the truncation operator set must hold one operator"; return TJSQ_ERR_ARG; }
  BasisView bs;
  read_basis(basis, bs);
  Ctx cx;
  Projector pj{pg, tops[0], cx};
  for (const Op& o : ops) pj.project(upd_block_ndia(o, bs, basis.length, flag, side, cx), false);
  if (site) pj.project(upd_site_ndia(st, bs, basis.length, flag, side), false);
  pg.x_elems = ops.empty() ? 0 : max_extent(*ops_in);
  pg.ops_elems = max_extent(trun);
  return finish(pg, shape, cx, msg);
}

namespace {

// an operator's sub-blocks as a contraction routine sees them, optionally
// after block_transfer_trans (block.f90:1676-1706): the transpose's quantum
// numbers (ket + the sub-block's shift, shift negated, rows and columns
// swapped); the data stay in place and are read transposed (t)
struct VSub { int num_up, num_down, down_dif, row_dim, sdim; int64_t off; double val; };
struct VOp { int up_dif = 0, down_dif = 0; bool t = false; std::vector<VSub> sub; };
VOp view_op(const Op& o, bool trans, int64_t base) {
  VOp v;
  v.up_dif = trans ? -o.up_dif : o.up_dif;
  v.down_dif = o.down_dif;
  v.t = trans;
  for (const OpSub& q : o.sub)
    v.sub.push_back(trans ? VSub{q.num_up + o.up_dif, q.num_down + q.down_dif, -q.down_dif, q.sdim, q.row_dim,
                                 base + q.off, 0.0}
                          : VSub{q.num_up, q.num_down, q.down_dif, q.row_dim, q.sdim, base + q.off, 0.0});
  return v;
}
VOp view_site(const Site& st, bool trans) {
  VOp v;
  v.up_dif = trans ? -st.up_dif : st.up_dif;
  v.down_dif = st.down_dif;
  for (size_t i = 0; i < st.sub.size(); ++i) {
    const OpSub& q = st.sub[i];
    v.sub.push_back(trans ? VSub{q.num_up + st.up_dif, q.num_down + q.down_dif, -q.down_dif, 1, 1, 0, st.val[i]}
                          : VSub{q.num_up, q.num_down, q.down_dif, 1, 1, 0, st.val[i]});
  }
  return v;
}
int first_vsub(const VOp& o, int up, int down, int dd) {
  for (size_t m = 0; m < o.sub.size(); ++m)
    if (o.sub[m].num_up == up && o.sub[m].num_down == down && o.sub[m].down_dif == dd) return (int)m;
  return -1;
}
// first sub-sector of the first sector (up, down) of b with the given block
// and site parts (the reference's nested searches with goto on a match);
// bl_up < -100000 leaves the block's particle label unchecked
const MapSub* first_sub(const BasisView& b, int up, int down, int bl_up, int bl_down, int st_up, int st_down) {
  for (const MapSec& m : b.sec) {
    if (m.new_up != up || m.new_down != down) continue;
    for (const MapSub& n : m.sub)
      if ((bl_up < -100000 || n.bl_up == bl_up) && n.bl_down == bl_down && n.st_up == st_up && n.st_down == st_down)
        return &n;
  }
  return nullptr;
}
double fsign(char flag, int n) { return flag == 'F' && n % 2 != 0 ? -1.0 : 1.0; }

}  // namespace

// sys_block_env_block_cor_ndia (measure.f90:5241-5480) for each environment
// block operator, sys_block_env_site_cor_ndia (:5482-5706) for the
// environment bond-site operator and sys_block_site_cor_ndia (:4939-5083)
// for the system bond-site operator, each against the one system block
// operator A.  Output copy q (environment block operators first, then the
// environment site, then the system site) holds Y_q with the routine's value
// = psi . Y_q.  The program's operator space is A's data followed by the
// environment operators' data; *_trans applies block_transfer_trans to those
// operands first, as the callers do for fermions.
int build_cor_ndia_program(const tjsq_opset& sys_op, const tjsq_opset* env_ops, int env_trans,
                           const tjsq_opset* env_site, int env_site_trans, const tjsq_opset* sys_site,
                           int sys_site_trans, const tjsq_basis& sys_basis, const tjsq_basis& env_basis, int nw,
                           const int* wsec, char flag, const Shape& shape, Program& pg, std::string& msg) {
  pg = Program();
  if (nw < 1 || !wsec || (flag != 'B' && flag != 'F')) { msg = "bad correlation arguments"; return TJSQ_ERR_ARG; }
  std::vector<Op> av, ev;
  int rc;
  if ((rc = read_ops(sys_op, av, msg))) return rc;
  if (sys_op.nops != 1) { msg = "one system block operator"; return TJSQ_ERR_ARG; }
  if (env_ops && env_ops->nops > 0 && (rc = read_ops(*env_ops, ev, msg))) return rc;
  Site est, sst;
  if (env_site && (rc = read_site(*env_site, est, msg))) return rc;
  if (sys_site && (rc = read_site(*sys_site, sst, msg))) return rc;
  const int64_t ebase = max_extent(sys_op);
  const VOp A = view_op(av[0], false, 0);
  std::vector<VOp> B;
  for (const Op& o : ev) B.push_back(view_op(o, env_trans != 0, ebase));
  const VOp ET = view_site(est, env_site_trans != 0), STv = view_site(sst, sys_site_trans != 0);
  BasisView sb, eb;
  read_basis(sys_basis, sb);
  read_basis(env_basis, eb);
  std::vector<WSec> W(nw);
  int64_t D = 0;
  for (int i = 0; i < nw; ++i) {
    const int* w = wsec + 6 * i;
    W[i] = {w[0], w[1], w[2], w[3], w[4], w[5], D};
    D += (int64_t)w[4] * w[5];
  }
  auto wave_of = [&](int su, int sd, int eu, int ed) {
    for (int m = 0; m < nw; ++m)
      if (W[m].sys_up == su && W[m].sys_down == sd && W[m].env_up == eu && W[m].env_down == ed) return m;
    return -1;
  };
  const int ncopy = (int)B.size() + (env_site ? 1 : 0) + (sys_site ? 1 : 0);
  // every region a term reads or writes lies inside its wavefunction sector
  bool bad = false;
  auto inside = [&](int w, int r0, int M, int c0, int N) {
    if (r0 < 0 || c0 < 0 || r0 + M > W[w].sys_dim || c0 + N > W[w].env_dim) bad = true;
    return !bad;
  };
  // one output block per (copy, sector, row origin, column origin): the
  // regions of one copy are the tiles of the fused sub-sector grid (sector
  // wide row strips for the system-site copy), each with one size
  std::map<std::tuple<int, int, int, int>, int> region;
  auto out_block = [&](int q, int w, int r0, int c0, int M, int N) {
    auto key = std::make_tuple(q, w, r0, c0);
    auto it = region.find(key);
    if (it != region.end()) {
      const OutBlock& b = pg.hp.out_blocks[it->second];
      if (b.M != M || b.N != N) bad = true;
      return it->second;
    }
    pg.hp.out_blocks.push_back({(int64_t)q * D + W[w].off + r0 + (int64_t)c0 * W[w].sys_dim,
                                std::max(1, W[w].sys_dim), M, N, (int)pg.hp.out_blocks.size()});
    const int id = (int)pg.hp.out_blocks.size() - 1;
    region.emplace(key, id);
    return id;
  };
  const int ud = A.up_dif, k1 = A.down_dif;
  for (int i = 0; i < nw; ++i) {
    const WSec& wi = W[i];
    if (wi.sys_dim == 0 || wi.env_dim == 0) continue;
    const int j = find_sec(sb, wi.sys_up, wi.sys_down), k = find_sec(eb, wi.env_up, wi.env_down);
    if (j < 0 || k < 0) continue;
    for (const MapSub& x : sb.sec[j].sub)
      for (const MapSub& y : eb.sec[k].sub) {
        const int64_t xoff = wi.off + x.spos + (int64_t)y.spos * wi.sys_dim;
        // (a) A (x) env block operator q
        for (size_t q = 0; q < B.size(); ++q)
          for (int dd = -k1; dd <= k1; dd += SU) {
            const int nbl_up = x.bl_up + ud, nbl_down = x.bl_down + dd;
            const int sid = first_vsub(A, x.bl_up, x.bl_down, dd);
            for (int edd = -k1; edd <= k1; edd += SU) {
              const int nebl_up = y.bl_up - ud, nebl_down = y.bl_down - edd;
              const int eid = first_vsub(B[q], nebl_up, nebl_down, edd);
              for (int nsd = std::abs(wi.sys_down - k1); nsd <= wi.sys_down + k1; nsd += SU) {
                const int nsu = wi.sys_up + ud, neu = wi.env_up - ud, ned = nsd;
                const MapSub* xp = first_sub(sb, nsu, nsd, nbl_up, nbl_down, x.st_up, x.st_down);
                const MapSub* yp = first_sub(eb, neu, ned, nebl_up, nebl_down, y.st_up, y.st_down);
                const int w = wave_of(nsu, nsd, neu, ned);
                if (sid < 0 || eid < 0 || !xp || !yp || w < 0) continue;
                double c1 = w6js(nsd, k1, wi.sys_down, x.bl_down, x.st_down, nbl_down);
                if (c1 == 0.0) continue;
                c1 *= w6js(ned, k1, wi.env_down, y.bl_down, y.st_down, nebl_down);
                if (c1 == 0.0) continue;
                c1 *= std::sqrt((1.0 + ned) * (1.0 + wi.env_down) / (1.0 + k1));
                if (((nbl_down + x.st_down + wi.sys_down + k1) / 2 + (nebl_down + y.st_down + wi.env_down + k1) / 2) %
                        2 != 0)
                  c1 = -c1;
                const VSub& as = A.sub[sid];
                const VSub& bs = B[q].sub[eid];
                if (as.row_dim != xp->sdim || as.sdim != x.sdim || bs.row_dim != y.sdim || bs.sdim != yp->sdim ||
                    !inside(i, x.spos, x.sdim, y.spos, y.sdim) || !inside(w, xp->spos, xp->sdim, yp->spos, yp->sdim)) {
                  msg = "operator sub-block disagrees with the fused bases";
                  return TJSQ_ERR_ARG;
                }
                Contrib c{};
                c.kind = C_BOTH;
                c.c = fsign(flag, wi.sys_down + y.st_down) * c1;
                c.a_off = as.off; c.a_ld = as.row_dim; c.a_t = 0;
                c.x_off = xoff; c.x_ld = wi.sys_dim; c.x_t = 0;
                c.b_off = bs.off; c.b_ld = B[q].t ? bs.sdim : bs.row_dim; c.b_t = B[q].t ? 1 : 0;
                c.M = xp->sdim; c.K1 = x.sdim; c.K2 = y.sdim; c.N = yp->sdim;
                c.out_block = out_block((int)q, w, xp->spos, yp->spos, c.M, c.N);
                if (c.M && c.N && c.K1 && c.K2) pg.hp.contrib.push_back(c);
              }
            }
          }
        // (b) A (x) the environment bond-site operator
        if (env_site) {
          const int q = (int)B.size();
          for (int dd = -k1; dd <= k1; dd += SU) {
            const int nbl_up = x.bl_up + ud, nbl_down = x.bl_down + dd;
            const int sid = first_vsub(A, x.bl_up, x.bl_down, dd);
            for (int sdd = -ET.down_dif; sdd <= ET.down_dif; sdd += SU) {
              const int nst_up = y.st_up - ud, nst_down = y.st_down - sdd;
              const int eid = first_vsub(ET, nst_up, nst_down, sdd);
              if (sid < 0 || eid < 0) continue;
              for (int nsd = std::abs(wi.sys_down - k1); nsd <= wi.sys_down + k1; nsd += SU) {
                const int nsu = wi.sys_up + ud, neu = wi.env_up - ud, ned = nsd;
                const MapSub* xp = first_sub(sb, nsu, nsd, nbl_up, nbl_down, x.st_up, x.st_down);
                const MapSub* yp = first_sub(eb, neu, ned, -1000000, y.bl_down, nst_up, nst_down);
                const int w = wave_of(nsu, nsd, neu, ned);
                if (!xp || !yp || w < 0) continue;
                double c1 = w6js(nsd, k1, wi.sys_down, x.bl_down, x.st_down, nbl_down);
                if (c1 == 0.0) continue;
                c1 *= w6js(ned, k1, wi.env_down, y.st_down, y.bl_down, nst_down);
                if (c1 == 0.0) continue;
                c1 *= std::sqrt((1.0 + ned) * (1.0 + wi.env_down) / (1.0 + k1));
                if (((nbl_down + x.st_down + wi.sys_down + k1) / 2 + (y.bl_down + y.st_down + ned + k1) / 2) % 2 != 0)
                  c1 = -c1;
                const VSub& as = A.sub[sid];
                if (as.row_dim != xp->sdim || as.sdim != x.sdim || yp->sdim != y.sdim ||
                    !inside(i, x.spos, x.sdim, y.spos, y.sdim) || !inside(w, xp->spos, xp->sdim, yp->spos, y.sdim)) {
                  msg = "operator sub-block disagrees with the fused bases";
                  return TJSQ_ERR_ARG;
                }
                Contrib c{};
                c.kind = C_LEFT;
                c.c = fsign(flag, wi.sys_down) * ET.sub[eid].val * c1;
                c.a_off = as.off; c.a_ld = as.row_dim; c.a_t = 0;
                c.x_off = xoff; c.x_ld = wi.sys_dim;
                c.M = xp->sdim; c.K1 = x.sdim; c.N = y.sdim;
                c.out_block = out_block(q, w, xp->spos, yp->spos, c.M, c.N);
                if (c.c != 0.0 && c.M && c.N && c.K1) pg.hp.contrib.push_back(c);
              }
            }
          }
        }
      }
    // (c) A (x) the system bond-site operator, inside the sector
    if (sys_site) {
      const int q = ncopy - 1;
      for (const MapSub& x : sb.sec[j].sub)
        for (int dd = -k1; dd <= k1; dd += SU)
          for (int sdd = -k1; sdd <= k1; sdd += SU) {
            const int nbl_up = x.bl_up + ud, nbl_down = x.bl_down + dd;
            const int nst_up = x.st_up - ud, nst_down = x.st_down - sdd;
            const int bid = first_vsub(A, x.bl_up, x.bl_down, dd);
            const int tid = first_vsub(STv, nst_up, nst_down, sdd);
            const MapSub* xp = nullptr;
            for (const MapSub& m : sb.sec[j].sub)
              if (m.bl_up == nbl_up && m.bl_down == nbl_down && m.st_up == nst_up && m.st_down == nst_down) {
                xp = &m;
                break;
              }
            if (bid < 0 || tid < 0 || !xp) continue;
            double c1 = w6js(nbl_down, x.bl_down, k1, x.st_down, nst_down, wi.sys_down) / std::sqrt(1.0 + k1);
            if (c1 == 0.0) continue;
            if (((x.bl_down + nst_down + k1 + wi.sys_down) / 2) % 2 != 0) c1 = -c1;
            const VSub& as = A.sub[bid];
            if (as.row_dim != xp->sdim || as.sdim != x.sdim || !inside(i, x.spos, x.sdim, 0, wi.env_dim) ||
                !inside(i, xp->spos, xp->sdim, 0, wi.env_dim)) {
              msg = "operator sub-block disagrees with the fused basis";
              return TJSQ_ERR_ARG;
            }
            Contrib c{};
            c.kind = C_LEFT;
            c.c = fsign(flag, x.bl_down) * STv.sub[tid].val * c1;
            c.a_off = as.off; c.a_ld = as.row_dim; c.a_t = 0;
            c.x_off = wi.off + x.spos; c.x_ld = wi.sys_dim;
            c.M = xp->sdim; c.K1 = x.sdim; c.N = wi.env_dim;
            c.out_block = out_block(q, i, xp->spos, 0, c.M, c.N);
            if (c.c != 0.0 && c.M && c.N && c.K1) pg.hp.contrib.push_back(c);
          }
    }
  }
  if (bad) { msg = "a correlation term falls outside its wavefunction sector"; return TJSQ_ERR_ARG; }
  pg.out.nops = ncopy;
  for (int q = 0; q < ncopy; ++q) pg.out.hdr.insert(pg.out.hdr.end(), {sys_basis.length, 0, 0, 0});
  pg.out.ndata = (int64_t)ncopy * D;
  pg.x_elems = D;
  pg.ops_elems = ebase + (env_ops && env_ops->nops > 0 ? max_extent(*env_ops) : 0);
  Ctx cx;
  return finish(pg, shape, cx, msg);
}

namespace {

// block_mul_block_ndia (block.f90:1956-2197) with Flag1 = Flag2 = 'N' and
// coef 1: the rank-dd coupling P = [O1 x O2]^dd of two embedded operators
// over one fused basis (new_oper%down_dif = dd, set by the caller there).
// Per basis sector i (ket, 2S = J), bra 2S nJ in |J - dd| .. J + dd and
// intermediate 2S m in |J - k2| .. J + k2,
//   P[i, nJ - J] += c O1[(m) -> nJ] O2[i -> m],
//   c = sum over mu of (nJ k1 m; -nJ nJ-mu mu) (m k2 J; -mu mu-J J) (-1)^((m-mu)/2)
//       (k1 k2 dd; nJ-mu mu-J J-nJ) (-1)^((k2-k1+nJ-J)/2) sqrt(dd+1)
//       / (nJ dd J; -nJ nJ-J J)
// (k1, k2 the ranks).  The product of a placement of O1 with an identity
// placement of O2 over the same sub-sector is one placement of O1's data;
// products of two data placements are not formed.
SymOp mul_ndia(const SymOp& o1, const SymOp& o2, int dd, const BasisView& bs, int len, Ctx& cx) {
  SymOp p;
  p.present = true; p.len = len; p.up_dif = o1.up_dif + o2.up_dif; p.down_dif = dd;
  const int k1 = o1.down_dif, k2 = o2.down_dif;
  auto find = [](const SymOp& o, int up, int down, int d) {
    for (size_t x = 0; x < o.sub.size(); ++x)
      if (o.sub[x].num_up == up && o.sub[x].num_down == down && o.sub[x].down_dif == d) return (int)x;
    return -1;
  };
  for (const MapSec& sec : bs.sec) {
    const int J = sec.new_down;
    for (int nJ = std::abs(J - dd); nJ <= J + dd; nJ += SU) {
      const int bsi = find_sec(bs, sec.new_up + p.up_dif, nJ);
      if (bsi < 0) continue;
      p.sub.push_back({sec.new_up, J, nJ - J, bs.sec[bsi].dim, sec.dim, {}});
      SymSub& d = p.sub.back();
      const int kq = nJ - J;
      for (int m = std::abs(J - k2); m <= J + k2; m += SU) {
        const int i2 = find(o2, sec.new_up, J, m - J);
        const int i1 = find(o1, sec.new_up + o2.up_dif, m, nJ - m);
        if (i1 < 0 || i2 < 0) continue;
        const double c11 = w3js(nJ, dd, J, -nJ, kq, J);
        if (c11 == 0.0) { cx.fail("vanishing normalising 3j symbol"); return p; }
        double c1 = 0.0;
        for (int mu = -m; mu <= m; mu += SU) {
          const int j5 = nJ - mu;
          double c12 = w3js(nJ, k1, m, -nJ, j5, mu);
          c12 = c12 * w3js(m, k2, J, -mu, mu - J, J) * par((m - mu) / 2);
          c12 = c12 * w3js(k1, k2, dd, j5, mu - J, -kq) * par((-k1 + k2 + kq) / 2) * std::sqrt(dd + 1.0);
          c1 = c1 + c12;
        }
        if (c1 == 0.0) continue;
        c1 = 1.0 * c1 / c11;
        for (const Place& a : o1.sub[i1].pl)
          for (const Place& b : o2.sub[i2].pl) {
            if (a.c0 + a.w <= b.r0 || b.r0 + b.h <= a.c0) continue;
            if (a.c0 != b.r0 || a.w != b.h) { cx.fail("operator product over misaligned placements"); return p; }
            if (a.src >= 0 && b.src >= 0) { cx.fail("operator product of two data blocks is not supported"); return p; }
            const double c = c1 * a.coef * b.coef;
            if (c == 0.0) continue;
            if (b.src < 0)
              d.pl.push_back({a.r0, b.c0, a.h, b.w, c, a.src, a.ld, a.t});
            else
              d.pl.push_back({a.r0, b.c0, a.h, b.w, c, b.src, b.ld, b.t});
          }
      }
    }
  }
  return p;
}

}  // namespace

// The same-side pair operators of Get_sys2b / Get_env2b (measure.f90:
// 2748-2819) at the block length where the second site joins: every
// operator of ops (rank k, block length L - 1) embedded into `basis`
// (update_block_ndia_*), the site operator embedded (update_site_ndia_*),
// their rank-0 coupling (block_mul_block_ndia), projected with T
// (update_trun_dia).  The outputs are diagonal operators at length L.
int build_pair_ndia_program(int side, char flag, const tjsq_opset& ops_in, const tjsq_opset& site,
                            const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                            std::string& msg) {
  pg = Program();
  if ((side != 0 && side != 1) || (flag != 'B' && flag != 'F')) { msg = "bad side or sign flag"; return TJSQ_ERR_ARG; }
  std::vector<Op> ops, tops;
  int rc;
  if ((rc = read_ops(ops_in, ops, msg))) return rc;
  Site st;
  if ((rc = read_site(site, st, msg))) return rc;
  if ((rc = read_ops(trun, tops, msg))) return rc;
  if (trun.nops != 1) { msg = "the truncation operator set must hold one operator"; return TJSQ_ERR_ARG; }
  BasisView bs;
  read_basis(basis, bs);
  Ctx cx;
  Projector pj{pg, tops[0], cx};
  const SymOp so = upd_site_ndia(st, bs, basis.length, flag, side);
  for (const Op& o : ops) {
    if (o.up_dif + st.up_dif != 0) { msg = "the pair operator must conserve the particle number"; return TJSQ_ERR_ARG; }
    pj.project(mul_ndia(upd_block_ndia(o, bs, basis.length, flag, side, cx), so, 0, bs, basis.length, cx), true);
  }
  pg.x_elems = max_extent(ops_in);
  pg.ops_elems = max_extent(trun);
  return finish(pg, shape, cx, msg);
}

// The general product of Get_sys_two / Get_sys_two1 / Get_SCOP_Operator_Env
// (measure.f90:3798-3889, 6211-6263): each operator of ops (block length
// L - 1) embedded into `basis` (update_block_ndia_*), the site operator
// embedded (update_site_ndia_*), their rank-`rank` coupling
// (block_mul_block_ndia), projected with T: update_trun_dia when the product
// is a scalar that conserves the particle label, update_trun_ndia otherwise.
int build_product_ndia_program(int side, char flag, int rank, const tjsq_opset& ops_in, const tjsq_opset& site,
                               const tjsq_basis& basis, const tjsq_opset& trun, const Shape& shape, Program& pg,
                               std::string& msg) {
  pg = Program();
  if ((side != 0 && side != 1) || (flag != 'B' && flag != 'F') || rank < 0 || rank > 4) {
    msg = "bad side, sign flag or rank";
    return TJSQ_ERR_ARG;
  }
  std::vector<Op> ops, tops;
  int rc;
  if ((rc = read_ops(ops_in, ops, msg))) return rc;
  Site st;
  if ((rc = read_site(site, st, msg))) return rc;
  if ((rc = read_ops(trun, tops, msg))) return rc;
  if (trun.nops != 1) { msg = "the truncation operator set must hold one operator"; return TJSQ_ERR_ARG; }
  BasisView bs;
  read_basis(basis, bs);
  Ctx cx;
  Projector pj{pg, tops[0], cx};
  const SymOp so = upd_site_ndia(st, bs, basis.length, flag, side);
  for (const Op& o : ops) {
    if ((o.down_dif + st.down_dif + rank) % 2 != 0 || rank > o.down_dif + st.down_dif ||
        rank < std::abs(o.down_dif - st.down_dif)) {
      msg = "the ranks of the factors cannot couple to the requested rank";
      return TJSQ_ERR_ARG;
    }
    const SymOp prod = mul_ndia(upd_block_ndia(o, bs, basis.length, flag, side, cx), so, rank, bs, basis.length, cx);
    pj.project(prod, rank == 0 && prod.up_dif == 0);
  }
  pg.x_elems = max_extent(ops_in);
  pg.ops_elems = max_extent(trun);
  return finish(pg, shape, cx, msg);
}

namespace {

// first sub-sector of the first sector (up, down) of b whose fields equal
// the given ones; a field given as ANY is not compared (the reference's
// searches compare three of the four block / site labels)
constexpr int ANY = -1000000;
const MapSub* find_sub(const BasisView& b, int up, int down, int bl_up, int bl_down, int st_up, int st_down) {
  for (const MapSec& m : b.sec) {
    if (m.new_up != up || m.new_down != down) continue;
    for (const MapSub& n : m.sub)
      if ((bl_up == ANY || n.bl_up == bl_up) && (bl_down == ANY || n.bl_down == bl_down) &&
          (st_up == ANY || n.st_up == st_up) && (st_down == ANY || n.st_down == st_down))
        return &n;
  }
  return nullptr;
}

}  // namespace

// The bond-site contractions of Get_SC_Order (measure.f90:3226-3376):
//   kind 0 sys_site_env_site_cor_ndia (:5709-5924)   sys_site x env_site
//   kind 1 sys_site_env_block_cor_ndia (:5926-6154)  sys_site x each env_ops operator
//   kind 2 env_block_site_cor_ndia (:5085-5238)      each env_ops operator x env_site
// Output copy q holds Y_q with the routine's value = psi . Y_q (one copy for
// kind 0, one per operator of env_ops otherwise).  *_trans applies
// block_transfer_trans to that operand first.  The program's operator space
// is env_ops' data (kinds 1, 2); x = psi, y cleared first.
int build_site_cor_program(int kind, const tjsq_opset* sys_site, int sys_site_trans, const tjsq_opset* env_ops,
                           int env_trans, const tjsq_opset* env_site, int env_site_trans,
                           const tjsq_basis& sys_basis, const tjsq_basis& env_basis, int nw, const int* wsec,
                           char flag, const Shape& shape, Program& pg, std::string& msg) {
  pg = Program();
  if (kind < 0 || kind > 2 || nw < 1 || !wsec || (flag != 'B' && flag != 'F')) {
    msg = "bad correlation arguments";
    return TJSQ_ERR_ARG;
  }
  const bool need_sys = kind <= 1, need_ops = kind >= 1, need_est = kind != 1;
  if ((need_sys && !sys_site) || (need_ops && (!env_ops || env_ops->nops < 1)) || (need_est && !env_site)) {
    msg = "missing operand for this contraction";
    return TJSQ_ERR_ARG;
  }
  int rc;
  Site sst, est;
  if (need_sys && (rc = read_site(*sys_site, sst, msg))) return rc;
  if (need_est && (rc = read_site(*env_site, est, msg))) return rc;
  std::vector<Op> ev;
  if (need_ops && (rc = read_ops(*env_ops, ev, msg))) return rc;
  std::vector<VOp> B;
  for (const Op& o : ev) B.push_back(view_op(o, env_trans != 0, 0));
  const VOp S = view_site(sst, sys_site_trans != 0), E = view_site(est, env_site_trans != 0);
  BasisView sb, eb;
  read_basis(sys_basis, sb);
  read_basis(env_basis, eb);
  std::vector<WSec> W(nw);
  int64_t D = 0;
  for (int i = 0; i < nw; ++i) {
    const int* w = wsec + 6 * i;
    if (w[4] < 0 || w[5] < 0) { msg = "negative sector dimension"; return TJSQ_ERR_ARG; }
    W[i] = {w[0], w[1], w[2], w[3], w[4], w[5], D};
    D += (int64_t)w[4] * w[5];
  }
  auto wave_of = [&](int su, int sd, int eu, int ed) {
    for (int m = 0; m < nw; ++m)
      if (W[m].sys_up == su && W[m].sys_down == sd && W[m].env_up == eu && W[m].env_down == ed) return m;
    return -1;
  };
  bool bad = false;
  auto inside = [&](int w, int r0, int M, int c0, int N) {
    if (r0 < 0 || c0 < 0 || r0 + M > W[w].sys_dim || c0 + N > W[w].env_dim) bad = true;
    return !bad;
  };
  std::map<std::tuple<int, int, int, int>, int> region;
  auto out_block = [&](int q, int w, int r0, int c0, int M, int N) {
    auto key = std::make_tuple(q, w, r0, c0);
    auto it = region.find(key);
    if (it != region.end()) {
      const OutBlock& b = pg.hp.out_blocks[it->second];
      if (b.M != M || b.N != N) bad = true;
      return it->second;
    }
    pg.hp.out_blocks.push_back({(int64_t)q * D + W[w].off + r0 + (int64_t)c0 * W[w].sys_dim,
                                std::max(1, W[w].sys_dim), M, N, (int)pg.hp.out_blocks.size()});
    const int id = (int)pg.hp.out_blocks.size() - 1;
    region.emplace(key, id);
    return id;
  };
  const int ncopy = kind == 0 ? 1 : (int)B.size();
  for (int i = 0; i < nw && !bad; ++i) {
    const WSec& wi = W[i];
    if (wi.sys_dim == 0 || wi.env_dim == 0) continue;
    const int j = find_sec(sb, wi.sys_up, wi.sys_down), k = find_sec(eb, wi.env_up, wi.env_down);
    if (k < 0 || (kind <= 1 && j < 0)) continue;
    if (kind == 2) {
      // env_block_site_cor_ndia: the environment block operator and site
      // operator on one sector; Y[:, new sub-sector] += c psi[:, x] B
      for (size_t q = 0; q < B.size(); ++q) {
        const int ud = B[q].up_dif, k1 = B[q].down_dif;
        for (const MapSub& x : eb.sec[k].sub)
          for (int bdd = -k1; bdd <= k1; bdd += SU) {
            const int nbu = x.bl_up - ud, nbd = x.bl_down - bdd;
            const int bid = first_vsub(B[q], nbu, nbd, bdd);
            for (int sdd = -k1; sdd <= k1; sdd += SU) {
              const int ntu = x.st_up + ud, ntd = x.st_down + sdd;
              const int tid = first_vsub(E, x.st_up, x.st_down, sdd);
              const MapSub* np = nullptr;
              for (const MapSub& n : eb.sec[k].sub)
                if (n.bl_up == nbu && n.bl_down == nbd && n.st_up == ntu && n.st_down == ntd) { np = &n; break; }
              if (bid < 0 || tid < 0 || !np) continue;
              const double sign = fsign(flag, 2 * wi.sys_down + x.st_down);
              double c1 = 1.0 / std::sqrt(1.0 + k1) * w6js(nbd, x.bl_down, k1, x.st_down, ntd, wi.env_down);
              c1 = c1 * par((x.bl_down + ntd + k1 + wi.env_down) / 2);
              if (c1 == 0.0) continue;
              const VSub& bs = B[q].sub[bid];
              if (bs.row_dim != x.sdim || bs.sdim != np->sdim || !inside(i, 0, wi.sys_dim, x.spos, x.sdim) ||
                  !inside(i, 0, wi.sys_dim, np->spos, np->sdim)) {
                msg = "operator sub-block disagrees with the fused basis";
                return TJSQ_ERR_ARG;
              }
              Contrib c{};
              c.kind = C_RIGHT;
              c.c = sign * E.sub[tid].val * c1;
              c.x_off = wi.off + (int64_t)x.spos * wi.sys_dim; c.x_ld = wi.sys_dim;
              c.b_off = bs.off; c.b_ld = B[q].t ? bs.sdim : bs.row_dim; c.b_t = B[q].t ? 1 : 0;
              c.M = wi.sys_dim; c.K2 = x.sdim; c.N = np->sdim;
              c.out_block = out_block((int)q, i, 0, np->spos, c.M, c.N);
              if (c.c != 0.0 && c.M && c.N && c.K2) pg.hp.contrib.push_back(c);
            }
          }
      }
      continue;
    }
    const int ud = S.up_dif, k1 = S.down_dif;
    for (const MapSub& x : sb.sec[j].sub)
      for (const MapSub& y : eb.sec[k].sub)
        for (int sdd = -k1; sdd <= k1; sdd += SU) {
          const int nsu_st = x.st_up + ud, nsd_st = x.st_down + sdd;
          const int sid = first_vsub(S, x.st_up, x.st_down, sdd);
          const int nq = kind == 0 ? 1 : (int)B.size();
          for (int q = 0; q < nq; ++q) {
            const VOp& R = kind == 0 ? E : B[q];
            // kind 0 walks st_down_dif2 over the system site's rank, kind 1
            // bl_down_dif over the environment block operator's rank
            const int rk = kind == 0 ? k1 : R.down_dif;
            for (int edd = -rk; edd <= rk; edd += SU) {
              int eid;
              int nebu = 0, nebd = 0, netu = 0, netd = 0;
              if (kind == 0) {
                netu = y.st_up - ud; netd = y.st_down - edd;
                eid = first_vsub(E, netu, netd, edd);
              } else {
                nebu = y.bl_up - ud; nebd = y.bl_down - edd;
                eid = first_vsub(R, nebu, nebd, edd);
              }
              if (sid < 0 || eid < 0) continue;
              for (int nsd = std::abs(wi.sys_down - k1); nsd <= wi.sys_down + k1; nsd += SU) {
                const int nsu = wi.sys_up + ud, neu = wi.env_up - ud, ned = nsd;
                const MapSub* xp = find_sub(sb, nsu, nsd, ANY, x.bl_down, nsu_st, nsd_st);
                const MapSub* yp = kind == 0 ? find_sub(eb, neu, ned, ANY, y.bl_down, netu, netd)
                                             : find_sub(eb, neu, ned, nebu, nebd, ANY, y.st_down);
                const int w = wave_of(nsu, nsd, neu, ned);
                if (!xp || !yp || w < 0) continue;
                double c1, sign;
                if (kind == 0) {
                  sign = fsign(flag, wi.sys_down + x.bl_down);
                  c1 = w6js(nsd, k1, wi.sys_down, x.st_down, x.bl_down, nsd_st);
                  if (c1 == 0.0) continue;
                  c1 *= w6js(ned, k1, wi.env_down, y.st_down, y.bl_down, netd);
                  if (c1 == 0.0) continue;
                  c1 *= std::sqrt((1.0 + ned) * (1.0 + wi.env_down) / (1.0 + k1));
                  if (((y.bl_down + y.st_down + ned + k1) / 2 + (x.bl_down + x.st_down + nsd + k1) / 2) % 2 != 0)
                    c1 = -c1;
                } else {
                  sign = fsign(flag, wi.sys_down + y.st_down + x.bl_down);
                  c1 = w6js(nsd, k1, wi.sys_down, x.st_down, x.bl_down, nsd_st);
                  if (c1 == 0.0) continue;
                  c1 *= w6js(ned, k1, wi.env_down, y.bl_down, y.st_down, nebd);
                  if (c1 == 0.0) continue;
                  c1 *= std::sqrt((1.0 + ned) * (wi.env_down + 1.0) / (1.0 + k1));
                  if (((nebd + y.st_down + wi.env_down + k1) / 2 + (x.bl_down + x.st_down + nsd + k1) / 2) % 2 != 0)
                    c1 = -c1;
                }
                Contrib c{};
                c.out_block = -1;
                if (kind == 0) {
                  if (xp->sdim != x.sdim || yp->sdim != y.sdim || !inside(i, x.spos, x.sdim, y.spos, y.sdim) ||
                      !inside(w, xp->spos, x.sdim, yp->spos, y.sdim)) {
                    msg = "site operator moves between sub-sectors of different sizes";
                    return TJSQ_ERR_ARG;
                  }
                  c.kind = C_AXPY;
                  c.c = sign * S.sub[sid].val * E.sub[eid].val * c1;
                  c.x_off = W[w].off + xp->spos + (int64_t)yp->spos * W[w].sys_dim; c.x_ld = W[w].sys_dim;
                  c.M = x.sdim; c.N = y.sdim;
                  c.out_block = out_block(0, i, x.spos, y.spos, c.M, c.N);
                  if (c.c != 0.0 && c.M && c.N) pg.hp.contrib.push_back(c);
                } else {
                  const VSub& bs = B[q].sub[eid];
                  if (xp->sdim != x.sdim || bs.row_dim != y.sdim || bs.sdim != yp->sdim ||
                      !inside(i, x.spos, x.sdim, y.spos, y.sdim) || !inside(w, xp->spos, x.sdim, yp->spos, yp->sdim)) {
                    msg = "operator sub-block disagrees with the fused bases";
                    return TJSQ_ERR_ARG;
                  }
                  // Y[x, y] += c psi_w[x', y'] B^T  (B maps y' -> y)
                  c.kind = C_RIGHT;
                  c.c = sign * S.sub[sid].val * c1;
                  c.x_off = W[w].off + xp->spos + (int64_t)yp->spos * W[w].sys_dim; c.x_ld = W[w].sys_dim;
                  c.b_off = bs.off; c.b_ld = B[q].t ? bs.sdim : bs.row_dim; c.b_t = B[q].t ? 0 : 1;
                  c.M = x.sdim; c.K2 = yp->sdim; c.N = y.sdim;
                  c.out_block = out_block(q, i, x.spos, y.spos, c.M, c.N);
                  if (c.c != 0.0 && c.M && c.N && c.K2) pg.hp.contrib.push_back(c);
                }
              }
            }
          }
        }
  }
  if (bad) { msg = "a correlation term falls outside its wavefunction sector"; return TJSQ_ERR_ARG; }
  pg.out.nops = ncopy;
  for (int q = 0; q < ncopy; ++q) pg.out.hdr.insert(pg.out.hdr.end(), {sys_basis.length, 0, 0, 0});
  pg.out.ndata = (int64_t)ncopy * D;
  pg.x_elems = D;
  pg.ops_elems = need_ops ? max_extent(*env_ops) : 0;
  Ctx cx;
  return finish(pg, shape, cx, msg);
}

// Get_density_matrix_sys (den_mat.f90:4-38): rho = psi psi^T per sector;
// Get_density_matrix_env (:44-78): rho = psi^T psi.
int build_density_program(int side, int len, int nw, const int* wsec, const Shape& shape, Program& pg,
                          std::string& msg) {
  pg = Program();
  if ((side != 0 && side != 1) || nw < 1 || !wsec) { msg = "bad density-matrix arguments"; return TJSQ_ERR_ARG; }
  int64_t xoff = 0;
  for (int i = 0; i < nw; ++i) {
    const int* w = wsec + 6 * i;
    const int ds = w[4], de = w[5];
    if (ds < 0 || de < 0) { msg = "negative sector dimension"; return TJSQ_ERR_ARG; }
    const int n = side == 0 ? ds : de;
    pg.out.sub.insert(pg.out.sub.end(), {side == 0 ? w[0] : w[2], side == 0 ? w[1] : w[3], 0, n, n});
    pg.hp.out_blocks.push_back({pg.out.ndata, std::max(1, n), n, n, i});
    pg.out.off.push_back(pg.out.ndata);
    pg.out.ndata += (int64_t)n * n;
    if (n > 0 && (side == 0 ? de : ds) > 0) {
      Contrib q{};
      q.kind = C_PROD;
      q.out_block = i;
      q.c = 1.0;
      q.M = n; q.N = n;
      q.a_space = SP_X; q.b_space = SP_X;
      q.a_off = xoff; q.a_ld = std::max(1, ds);
      q.b_off = xoff; q.b_ld = std::max(1, ds);
      if (side == 0) { q.a_t = 0; q.b_t = 1; q.K1 = de; }
      else { q.a_t = 1; q.b_t = 0; q.K1 = ds; }
      pg.hp.contrib.push_back(q);
    }
    xoff += (int64_t)ds * de;
  }
  pg.out.nops = 1;
  pg.out.hdr = {len, 0, 0, nw};
  pg.x_elems = xoff;
  pg.ops_elems = 0;
  Ctx cx;
  return finish(pg, shape, cx, msg);
}

// --------------------------------------------------------------------------
// wave_trans_LR / wave_trans_RL (dmrg_sub.f90:1093-1384)
// --------------------------------------------------------------------------
namespace {

// (-1)**n of an integer n, as Fortran's integer power gives it
double neg_pow(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

// recoupling coefficient of wave_trans_LR (dmrg_sub.f90:1207-1214):
// j12 = 2S of the old system sector, j3 = the moving site's 2S, j123 = the
// new system sector's 2S, j4 = the new environment sector's 2S, j34 = the
// old environment sector's 2S
double coef_lr(int j12, int j3, int j123, int j4, int j34) {
  double c = neg_pow((j12 + j3 + j123) / 2) / std::sqrt((1.0 + j12) * (1.0 + j123));
  if (c != 0.0) {
    c = c * std::sqrt((1.0 + j123) * (j34 + 1.0));
    if (((j12 + j3 + j4) / 2) % 2 != 0) c = -c;
    if (((j3 + j4 - j34) / 2) % 2 != 0) c = -c;
  }
  return c;
}

// wave_trans_RL (dmrg_sub.f90:1354-1361): j12 = old environment 2S, j3 =
// site 2S, j4 = new system 2S, j123 = new environment 2S, j34 = old system 2S
double coef_rl(int j12, int j3, int j4, int j123, int j34) {
  double c = neg_pow((j34 + j4 + j3) / 2) / std::sqrt((1.0 + j12) * (1.0 + j4));
  if (c != 0.0) {
    c = c * std::sqrt((1.0 + j123) * (j34 + 1.0));
    if (((j4 + j3 + j12) / 2) % 2 != 0) c = -c;
    if (((j12 + j3 - j123) / 2) % 2 != 0) c = -c;
  }
  return c;
}

int first_sub(const Op& o, int up, int down) {
  for (size_t k = 0; k < o.sub.size(); ++k)
    if (o.sub[k].num_up == up && o.sub[k].num_down == down) return (int)k;
  return -1;
}

// first sector of b with quantum numbers (up, down) holding a sub-block
// (bl_up, bl_down) x (st_up, st_down) (the reference's nested search with
// goto on the first match)
bool find_fused(const BasisView& b, int up, int down, int bl_up, int bl_down, int st_up, int st_down, int& y,
                int& z) {
  for (size_t s = 0; s < b.sec.size(); ++s) {
    if (b.sec[s].new_up != up || b.sec[s].new_down != down) continue;
    for (size_t q = 0; q < b.sec[s].sub.size(); ++q) {
      const MapSub& m = b.sec[s].sub[q];
      if (m.bl_up == bl_up && m.bl_down == bl_down && m.st_up == st_up && m.st_down == st_down) {
        y = (int)s;
        z = (int)q;
        return true;
      }
    }
  }
  return false;
}

int first_sector(const BasisView& b, int up, int down) {
  for (size_t s = 0; s < b.sec.size(); ++s)
    if (b.sec[s].new_up == up && b.sec[s].new_down == down) return (int)s;
  return -1;
}

}  // namespace

int build_wave_transform_program(int dir, const tjsq_basis& sys_bs, const tjsq_basis& env_bs,
                                 const tjsq_opset& sys_trun, const tjsq_opset& env_trun, int nw_old,
                                 const int* wsec_old, int nw_new, const int* wsec_new, const Shape& shape,
                                 Program& pg, std::string& msg) {
  pg = Program();
  if (dir != 0 && dir != 1) { msg = "direction must be 0 (left to right) or 1 (right to left)"; return TJSQ_ERR_ARG; }
  if (nw_old < 0 || nw_new < 1 || (nw_old > 0 && !wsec_old) || !wsec_new) {
    msg = "bad wavefunction sector lists";
    return TJSQ_ERR_ARG;
  }
  std::vector<Op> st, et;
  int rc = read_ops(sys_trun, st, msg);
  if (rc) return rc;
  if ((rc = read_ops(env_trun, et, msg))) return rc;
  if (sys_trun.nops != 1 || env_trun.nops != 1) { msg = "each truncation operator set must hold one operator"; return TJSQ_ERR_ARG; }
  const Op& S = st[0];
  const Op& E = et[0];
  BasisView SB, EB;
  read_basis(sys_bs, SB);
  read_basis(env_bs, EB);
  auto sectors = [&](int n, const int* w, std::vector<WSec>& out, int64_t& total) -> bool {
    out.resize(n);
    total = 0;
    for (int i = 0; i < n; ++i) {
      const int* q = w + 6 * i;
      if (q[4] < 0 || q[5] < 0) return false;
      out[i] = {q[0], q[1], q[2], q[3], q[4], q[5], total};
      total += (int64_t)q[4] * q[5];
    }
    return true;
  };
  std::vector<WSec> ow, nwv;
  int64_t dold = 0, dnew = 0;
  if (!sectors(nw_old, wsec_old, ow, dold) || !sectors(nw_new, wsec_new, nwv, dnew)) {
    msg = "negative sector dimension";
    return TJSQ_ERR_ARG;
  }
  const int64_t eoff = sys_trun.ndata;   // env_trun's data follows sys_trun's
  pg.x_elems = dold;
  pg.ops_elems = sys_trun.ndata + env_trun.ndata;
  pg.normalize_out = true;
  // output: one sub-block per new sector; work blocks are the fused
  // sub-blocks along the side that gained the site (LR: system rows, RL:
  // environment columns), which tile each sector exactly
  const BasisView& FB = dir == 0 ? SB : EB;
  std::vector<std::vector<int>> ob_of(nw_new);
  for (int w = 0; w < nw_new; ++w) {
    const WSec& n = nwv[w];
    pg.out.sub.insert(pg.out.sub.end(), {n.sys_up, n.sys_down, 0, n.sys_dim, n.env_dim});
    pg.out.off.push_back(n.off);
    const int fs = dir == 0 ? first_sector(FB, n.sys_up, n.sys_down) : first_sector(FB, n.env_up, n.env_down);
    const int extent = dir == 0 ? n.sys_dim : n.env_dim;
    int covered = 0;
    if (fs >= 0)
      for (const MapSub& m : FB.sec[fs].sub) {
        if (m.spos != covered || m.sdim < 0) { covered = -1; break; }
        covered += m.sdim;
      }
    if (fs < 0 || covered != extent) {
      msg = "a new wavefunction sector disagrees with its fused basis sector";
      return TJSQ_ERR_ARG;
    }
    for (size_t z = 0; z < FB.sec[fs].sub.size(); ++z) {
      const MapSub& m = FB.sec[fs].sub[z];
      ob_of[w].push_back((int)pg.hp.out_blocks.size());
      if (dir == 0)
        pg.hp.out_blocks.push_back({n.off + m.spos, std::max(1, n.sys_dim), m.sdim, n.env_dim, w});
      else
        pg.hp.out_blocks.push_back({n.off + (int64_t)m.spos * n.sys_dim, std::max(1, n.sys_dim), n.sys_dim, m.sdim, w});
    }
  }
  pg.out.nops = 1;
  pg.out.hdr = {0, 0, 0, nw_new};
  pg.out.ndata = dnew;
  auto find_new = [&](int su, int sd, int eu, int ed) -> int {
    for (int w = 0; w < nw_new; ++w)
      if (nwv[w].sys_up == su && nwv[w].sys_down == sd && nwv[w].env_up == eu && nwv[w].env_down == ed) return w;
    return -1;
  };
  const int SUs = 2;   // spin step (common.f90:11)
  for (int i = 0; i < nw_old; ++i) {
    const WSec& o = ow[i];
    if (dir == 0) {
      const int sid = first_sub(S, o.sys_up, o.sys_down);
      if (sid < 0) continue;
      for (const MapSec& k : EB.sec) {
        if (k.new_up != o.env_up || k.new_down != o.env_down) continue;
        for (const MapSub& x : k.sub) {
          const int eid = first_sub(E, x.bl_up, x.bl_down);
          for (int nsd = std::abs(o.sys_down - x.st_down); nsd <= o.sys_down + x.st_down; nsd += SUs) {
            const int nsu = o.sys_up + x.st_up;
            int y = -1, z = -1;
            const bool fused = find_fused(SB, nsu, nsd, o.sys_up, o.sys_down, x.st_up, x.st_down, y, z);
            const int w = find_new(nsu, nsd, x.bl_up, x.bl_down);
            if (eid < 0 || !fused || w < 0) continue;
            const double c = coef_lr(o.sys_down, x.st_down, nsd, x.bl_down, o.env_down);
            if (c == 0.0) continue;
            const OpSub& A = S.sub[sid];
            const OpSub& B = E.sub[eid];
            const MapSub& f = SB.sec[y].sub[z];
            if (A.row_dim != o.sys_dim || A.sdim != f.sdim || B.row_dim != nwv[w].env_dim || B.sdim != x.sdim ||
                x.spos + x.sdim > o.env_dim) {
              msg = "truncation operator dimensions disagree with the wavefunction sectors";
              return TJSQ_ERR_ARG;
            }
            if (f.sdim == 0 || o.sys_dim == 0 || x.sdim == 0 || nwv[w].env_dim == 0) continue;
            Contrib q{};
            q.kind = C_BOTH;
            q.out_block = ob_of[w][z];
            q.c = c;
            q.M = f.sdim; q.K1 = o.sys_dim; q.K2 = x.sdim; q.N = nwv[w].env_dim;
            q.a_off = A.off; q.a_ld = std::max(1, A.row_dim); q.a_t = 1;
            q.x_off = o.off + (int64_t)x.spos * o.sys_dim; q.x_ld = std::max(1, o.sys_dim); q.x_t = 0;
            q.b_off = eoff + B.off; q.b_ld = std::max(1, B.row_dim); q.b_t = 1;
            pg.hp.contrib.push_back(q);
          }
        }
      }
    } else {
      const int eid = first_sub(E, o.env_up, o.env_down);
      if (eid < 0) continue;
      for (const MapSec& k : SB.sec) {
        if (k.new_up != o.sys_up || k.new_down != o.sys_down) continue;
        for (const MapSub& x : k.sub) {
          for (int ned = std::abs(o.env_down - x.st_down); ned <= o.env_down + x.st_down; ned += SUs) {
            const int neu = o.env_up + x.st_up;
            const int sid = first_sub(S, x.bl_up, x.bl_down);
            int y = -1, z = -1;
            const bool fused = find_fused(EB, neu, ned, o.env_up, o.env_down, x.st_up, x.st_down, y, z);
            const int w = find_new(x.bl_up, x.bl_down, neu, ned);
            if (sid < 0 || !fused || w < 0) continue;
            const double c = coef_rl(o.env_down, x.st_down, x.bl_down, ned, o.sys_down);
            if (c == 0.0) continue;
            const OpSub& A = S.sub[sid];
            const OpSub& B = E.sub[eid];
            const MapSub& f = EB.sec[y].sub[z];
            if (A.row_dim != nwv[w].sys_dim || A.sdim != x.sdim || B.row_dim != o.env_dim || B.sdim != f.sdim ||
                x.spos + x.sdim > o.sys_dim) {
              msg = "truncation operator dimensions disagree with the wavefunction sectors";
              return TJSQ_ERR_ARG;
            }
            if (nwv[w].sys_dim == 0 || x.sdim == 0 || o.env_dim == 0 || f.sdim == 0) continue;
            Contrib q{};
            q.kind = C_BOTH;
            q.out_block = ob_of[w][z];
            q.c = c;
            q.M = nwv[w].sys_dim; q.K1 = x.sdim; q.K2 = o.env_dim; q.N = f.sdim;
            q.a_off = A.off; q.a_ld = std::max(1, A.row_dim); q.a_t = 0;
            q.x_off = o.off + x.spos; q.x_ld = std::max(1, o.sys_dim); q.x_t = 0;
            q.b_off = eoff + B.off; q.b_ld = std::max(1, B.row_dim); q.b_t = 0;
            pg.hp.contrib.push_back(q);
          }
        }
      }
    }
  }
  Ctx cx;
  return finish(pg, shape, cx, msg);
}

// --------------------------------------------------------------------------
// Get_truncation_operator, host part (den_mat.f90:141-383)
// --------------------------------------------------------------------------
void quicksort_double(double* a, int* idx, int lower, int upper) {
  // a[i - 1] holds element i
  int i = lower, j = upper;
  const int index = (int)((lower + upper) / 2.0);
  const double mn = a[index - 1];
  while (i <= j) {
    while (a[i - 1] > mn) ++i;
    while (a[j - 1] < mn) --j;
    if (i <= j) {
      std::swap(a[i - 1], a[j - 1]);
      std::swap(idx[i - 1], idx[j - 1]);
      ++i;
      --j;
    }
  }
  if (lower < j) quicksort_double(a, idx, lower, j);
  if (i < upper) quicksort_double(a, idx, i, upper);
}

int select_truncation(int nsec, const int* den_sub, const double* eigval, int kept_min, int kept_max,
                      double trun_err, TruncSelection& sel, std::string& msg) {
  sel = TruncSelection();
  if (nsec < 0 || (nsec > 0 && (!den_sub || !eigval))) { msg = "bad truncation arguments"; return TJSQ_ERR_ARG; }
  int dim = 0;
  for (int k = 0; k < nsec; ++k) {
    if (den_sub[5 * k + 3] != den_sub[5 * k + 4]) { msg = "density-matrix sector is not square"; return TJSQ_ERR_ARG; }
    dim += den_sub[5 * k + 4];
  }
  std::vector<double> ev(eigval, eigval + dim);
  std::vector<int> idx(dim), eigs(dim), sec_of(dim);
  for (int k = 0, p = 0; k < nsec; ++k)
    for (int i = 0; i < den_sub[5 * k + 4]; ++i, ++p) {
      idx[p] = p + 1;
      eigs[p] = den_sub[5 * k + 1] + 1;   // multiplicity 2S + 1
      sec_of[p] = k;
    }
  double tot_pm = 0.0;
  for (int i = 0; i < dim; ++i) tot_pm = tot_pm + ev[i] * eigs[i];
  double entropy1 = 0.0;
  for (int i = 0; i < dim; ++i)
    if (ev[i] >= 1.0E-15) {
      const double v = ev[i] / eigs[i];
      entropy1 = entropy1 - v * std::log(v) * eigs[i];
    }
  for (int i = 0; i < dim; ++i) ev[i] = ev[i] / tot_pm;
  double entropy2 = 0.0;
  for (int i = 0; i < dim; ++i)
    if (ev[i] >= 1.0E-15) entropy2 = entropy2 - ev[i] * std::log(ev[i]) * eigs[i];
  for (int i = 0; i < dim; ++i) ev[i] = ev[i] * tot_pm;
  if (dim > 0) quicksort_double(ev.data(), idx.data(), 1, dim);
  tot_pm = 0.0;
  for (int i = 0; i < dim; ++i) tot_pm = tot_pm + ev[i];
  for (int i = 0; i < dim; ++i) ev[i] = ev[i] / tot_pm;
  int num = 0;
  double err = 0.0;
  if (dim <= kept_min) {
    num = dim;
  } else {
    for (int i = kept_min + 1; i <= dim; ++i) err = err + ev[i - 1];
    num = kept_min;
    // the count stops at dim: the reference would read past the spectrum
    while (err > trun_err && num < dim) {
      ++num;
      err = err - ev[num - 1];
      if (err <= trun_err || num >= kept_max) break;
    }
  }
  double ent = 0.0;
  for (int i = 0; i < dim; ++i)
    if (ev[i] >= 1.0E-15) ent = ent - ev[i] * std::log(ev[i]);
  sel.counts.assign(nsec, 0);
  for (int k = 0; k < num; ++k) sel.counts[sec_of[idx[k] - 1]]++;
  sel.num_state = num;
  sel.dim = dim;
  sel.trunc_err = err;
  sel.entropy = ent;
  sel.entropy1 = entropy1;
  sel.entropy2 = entropy2;
  sel.sorted = ev;
  sel.order = idx;
  return TJSQ_OK;
}

}  // namespace tjsq
