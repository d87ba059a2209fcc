// tjsq_packed.cpp -- packed-operand task tables of the BF16x3 H_eff apply
// (layout and tiling described in tjsq_packed.hpp).
#include "tjsq_packed.hpp"

#include <algorithm>
#include <map>
#include <string>
#include <tuple>

namespace tjsq {

namespace {

inline int pad16(int n) { return (n + 15) & ~15; }

// registry of packed copies: one per (space, source offset, orientation)
struct Copies {
  std::map<std::tuple<int64_t, int>, int64_t> at;
  std::vector<PackJob>* jobs;
  int64_t* bytes;
  bool bad = false;
  std::map<std::tuple<int64_t, int>, size_t> job_of;
  // byte offset of the copy of the R x C source at src_off (leading dim ld)
  // with lines along rows (trans = 1) or along columns (trans = 0)
  int64_t get(int64_t src_off, int ld, int R, int C, int trans, int32_t& stride) {
    const int lines = trans ? pad16(R) : pad16(C), kpad = trans ? pad16(C) : pad16(R);
    stride = kpad * 4;
    auto key = std::make_tuple(src_off, trans);
    auto it = at.find(key);
    if (it != at.end()) {
      const PackJob& o = (*jobs)[job_of.at(key)];
      if (o.rows != R || o.cols != C || o.src_ld != ld) bad = true;
      return it->second;
    }
    PackJob j{src_off, *bytes, ld, R, C, trans, lines, kpad};
    job_of.emplace(key, jobs->size());
    jobs->push_back(j);
    at.emplace(key, *bytes);
    *bytes += (int64_t)lines * kpad * 4;
    return j.dst_off;
  }
};

// an operand op(M) (stored R x C column-major) read as lines: as the left
// operand, one line per row of op(M) with k along its columns; as the right
// operand, one line per column of op(M) with k along its rows
struct Line { int64_t off; int32_t stride; };

Line left_operand(Copies& cp, int64_t off, int ld, int t, int rows_op, int cols_op) {
  const int R = t ? cols_op : rows_op, C = t ? rows_op : cols_op;
  Line l;
  l.off = cp.get(off, ld, R, C, t ? 0 : 1, l.stride);
  return l;
}
Line right_operand(Copies& cp, int64_t off, int ld, int t, int rows_op, int cols_op) {
  const int R = t ? cols_op : rows_op, C = t ? rows_op : cols_op;
  Line l;
  l.off = cp.get(off, ld, R, C, t ? 1 : 0, l.stride);
  return l;
}

// Tiles of one locality group on one XCD: groups onto the least loaded
// of NX bins, largest first; within a bin the groups run in that order,
// their tiles in table order.  Workgroup g runs on XCD g mod NX, so bin x's
// q-th tile takes index q NX + x; bins are padded with empty tiles (no
// blocks, no segments: nothing read or written).
void xcd_bins(std::vector<PTile>& tiles, const std::vector<int>& group, const std::vector<int64_t>& cost, int NX,
              int out_kind, int unit = 1) {
  // items of `unit` consecutive tiles (a workgroup's), group (0 .. G - 1)
  // and cost per item
  const size_t n = tiles.size() / unit;
  int G = 0;
  for (size_t t = 0; t < n; ++t) G = std::max(G, group[t] + 1);
  std::vector<int64_t> gcost(G, 0);
  std::vector<char> used(G, 0);
  for (size_t t = 0; t < n; ++t) {
    gcost[group[t]] += cost[t];
    used[group[t]] = 1;
  }
  std::vector<int> order;
  for (int g = 0; g < G; ++g)
    if (used[g]) order.push_back(g);
  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return gcost[a] != gcost[b] ? gcost[a] > gcost[b] : a < b; });
  std::vector<int> gbin(G, 0), grank(G, 0);
  std::vector<int64_t> load(NX, 0);
  for (size_t r = 0; r < order.size(); ++r) {
    const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    gbin[order[r]] = b;
    grank[order[r]] = (int)r;
    load[b] += gcost[order[r]];
  }
  // items by group rank (a counting sort: stable, table order within a group)
  std::vector<int> start(order.size() + 1, 0);
  for (size_t t = 0; t < n; ++t) ++start[grank[group[t]] + 1];
  for (size_t r = 0; r < order.size(); ++r) start[r + 1] += start[r];
  std::vector<int> idx(n);
  for (size_t t = 0; t < n; ++t) idx[start[grank[group[t]]]++] = (int)t;
  std::vector<std::vector<int>> bins(NX);
  for (int t : idx) bins[gbin[group[t]]].push_back(t);
  size_t maxlen = 0;
  for (auto& b : bins) maxlen = std::max(maxlen, b.size());
  PTile null{};
  null.out_kind = out_kind;
  null.slot = -1;
  std::vector<PTile> sorted(maxlen * NX * unit, null);
  for (int x = 0; x < NX; ++x)
    for (size_t q = 0; q < bins[x].size(); ++q)
      for (int u = 0; u < unit; ++u) sorted[(q * NX + x) * unit + u] = tiles[bins[x][q] * unit + u];
  tiles.swap(sorted);
}

// work of a tile: K summed over its segments (an axpy counts 1) times its blocks
std::vector<int64_t> tile_costs(const std::vector<PTile>& tiles, const std::vector<PSeg>& segs) {
  std::vector<int64_t> cost(tiles.size(), 0);
  for (size_t t = 0; t < tiles.size(); ++t) {
    const PTile& tl = tiles[t];
    for (int q = tl.seg_begin; q < tl.seg_end; ++q) cost[t] += segs[q].K > 0 ? segs[q].K : 1;
    cost[t] = std::max<int64_t>(1, cost[t]) * tl.mb * tl.nb;
  }
  return cost;
}

}  // namespace

int compile_packed(const HostPlan& P, const PackShape& shape, PackedPlan& out, std::string& msg) {
  out = PackedPlan();
  if (shape.k_split < 32 || shape.k_split % 32 || shape.item_chunks < 1) {
    msg = "packed layout: k_split must be a positive multiple of 32";
    return TJSQ_ERR_ARG;
  }
  for (const Contrib& c : P.contrib)
    if (c.kind == C_PROD || c.x_t) {
      msg = "packed layout: contribution kind not covered";
      return TJSQ_ERR_ARG;
    }
  Copies ops;
  ops.jobs = &out.ops_jobs;
  ops.bytes = &out.ops_bytes;
  Copies xs;
  xs.jobs = &out.x_jobs;
  xs.bytes = &out.x_bytes;

  // ---- stage 1: T = op(A) op(X), computed as T^T = op(X)^T op(A)^T -------
  // T (M x K2) is stored as pad16(M) lines of pad16(K2): the left operand of
  // its stage-2 segment.  Sharing as in compile_program: contributions into
  // one output block with the same right operand are factored into one
  // multi-segment product; other products are shared by (A, op, X, M, K2).
  struct TRef { int64_t off; int32_t stride; };
  std::vector<TRef> t_of(P.contrib.size(), TRef{-1, 0});
  std::vector<int> lead(P.contrib.size(), -1);
  auto new_t = [&](int M, int K2) {
    TRef t{out.t_bytes, pad16(K2) * 4};
    out.t_bytes += (int64_t)pad16(M) * pad16(K2) * 4;
    return t;
  };
  auto stage1_segment = [&](const Contrib& c, double coef) {
    // left: lines over the columns of op(X) (k = its rows); right: lines over
    // the rows of op(A) (k = its columns)
    const Line lx = right_operand(xs, c.x_off, c.x_ld, 0, c.K1, c.K2);
    const Line la = left_operand(ops, c.a_off, c.a_ld, c.a_t, c.M, c.K1);
    PSeg s{};
    s.l_off = lx.off; s.l_stride = lx.stride;
    s.r_off = la.off; s.r_stride = la.stride;
    s.spaces = PK_X | (PK_OPS << 8);
    s.K = pad16(c.K1);
    s.c = coef;
    out.seg1.push_back(s);
  };
  auto stage1_tiles = [&](const Contrib& c, const TRef& t, int sb, int se) {
    const int U = pad16(c.K2), V = pad16(c.M);
    for (int u0 = 0; u0 < U; u0 += 64)
      for (int v0 = 0; v0 < V; v0 += 64) {
        PTile tl{};
        tl.out_off = t.off; tl.out_ld = t.stride;
        tl.m0 = u0; tl.n0 = v0;
        tl.mb = std::min(64, U - u0) / 16; tl.nb = std::min(64, V - v0) / 16;
        tl.mlim = U; tl.nlim = V;
        tl.seg_begin = sb; tl.seg_end = se;
        tl.out_kind = PK_OUT_T; tl.slot = -1;
        out.tile1.push_back(tl);
      }
  };
  {
    std::map<std::tuple<int, int64_t, int, int, int, int>, std::vector<int>> groups;
    for (size_t q = 0; q < P.contrib.size(); ++q) {
      const Contrib& c = P.contrib[q];
      if (c.kind == C_BOTH) groups[std::make_tuple(c.out_block, c.b_off, (int)c.b_t, c.b_ld, c.K2, c.N)].push_back((int)q);
    }
    for (auto& kv : groups) {
      const std::vector<int>& g = kv.second;
      if (g.size() < 2) continue;
      const Contrib& c0 = P.contrib[g[0]];
      const TRef t = new_t(c0.M, c0.K2);
      const int sb = (int)out.seg1.size();
      for (int q : g) {
        lead[q] = g[0];
        stage1_segment(P.contrib[q], P.contrib[q].c);
      }
      t_of[g[0]] = t;
      stage1_tiles(c0, t, sb, (int)out.seg1.size());
    }
  }
  {
    std::map<std::tuple<int64_t, int, int64_t, int, int>, TRef> tmap;
    for (size_t q = 0; q < P.contrib.size(); ++q) {
      const Contrib& c = P.contrib[q];
      if (c.kind != C_BOTH || lead[q] >= 0) continue;
      auto key = std::make_tuple(c.a_off, (int)c.a_t, c.x_off, c.M, c.K2);
      auto it = tmap.find(key);
      if (it != tmap.end()) {
        t_of[q] = it->second;
        continue;
      }
      const TRef t = new_t(c.M, c.K2);
      tmap.emplace(key, t);
      t_of[q] = t;
      const int sb = (int)out.seg1.size();
      stage1_segment(c, 1.0);
      stage1_tiles(c, t, sb, sb + 1);
    }
  }

  if (shape.order1 && !out.tile1.empty()) {
    // locality order: x block of the first segment (by its offset in x: sys
    // sector outer, env sector inner), then the operator copy
    std::map<int64_t, int64_t> x_src;
    for (const PackJob& j : out.x_jobs) x_src.emplace(j.dst_off, j.src_off);
    const size_t n = out.tile1.size();
    std::vector<int> order(n);
    std::vector<std::tuple<int64_t, int64_t, int, int>> key(n);
    for (size_t t = 0; t < n; ++t) {
      const PTile& tl = out.tile1[t];
      const PSeg& s0 = out.seg1[tl.seg_begin];
      key[t] = std::make_tuple(x_src.at(s0.l_off), s0.r_off, tl.n0, tl.m0);
      order[t] = (int)t;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<PTile> sorted(n);
    for (size_t i = 0; i < n; ++i) sorted[i] = out.tile1[order[i]];
    out.tile1.swap(sorted);
    const int unit = shape.quad1 ? 4 : 1;
    if (shape.quad1) {
      // quads: tiles with the same x rows (m0, mb) and the same x segments
      // (copy offset and K, segment by segment), in fours, in table order
      std::map<std::vector<int64_t>, int> gid;
      std::vector<std::vector<int>> groups;
      for (size_t t = 0; t < n; ++t) {
        const PTile& tl = out.tile1[t];
        std::vector<int64_t> key{tl.m0, tl.mb};
        for (int q = tl.seg_begin; q < tl.seg_end; ++q) {
          key.push_back(out.seg1[q].l_off);
          key.push_back(out.seg1[q].K);
        }
        auto it = gid.emplace(key, (int)groups.size());
        if (it.second) groups.emplace_back();
        groups[it.first->second].push_back((int)t);
      }
      PTile null{};
      null.out_kind = PK_OUT_T;
      null.slot = -1;
      std::vector<PTile> quads;
      for (const std::vector<int>& g : groups)
        for (size_t i = 0; i < g.size(); i += 4)
          for (size_t u = 0; u < 4; ++u) quads.push_back(i + u < g.size() ? out.tile1[g[i + u]] : null);
      out.tile1.swap(quads);
      out.quad1 = true;
    }
    if (shape.xcds1 > 1) {
      // locality group: the operator copy (right operand) of the tile (of a
      // quad's lead tile), so that one XCD reads each operator block for
      // every x block it meets
      const size_t nu = out.tile1.size() / unit;
      std::map<int64_t, int> gid;
      std::vector<int> g1(nu);
      std::vector<int64_t> c1(nu, 0);
      const std::vector<int64_t> ct = tile_costs(out.tile1, out.seg1);
      for (size_t t = 0; t < nu; ++t) {
        g1[t] = gid.emplace(out.seg1[out.tile1[t * unit].seg_begin].r_off, (int)gid.size()).first->second;
        for (int u = 0; u < unit; ++u) c1[t] += out.tile1[t * unit + u].mb > 0 ? ct[t * unit + u] : 0;
      }
      xcd_bins(out.tile1, g1, c1, shape.xcds1, PK_OUT_T, unit);
    }
  }

  // ---- stage 2: per output block, axpy segments first, then GEMM pieces --
  std::vector<int> order(P.contrib.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = (int)q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return P.contrib[a].out_block < P.contrib[b].out_block; });
  std::vector<std::vector<PSeg>> per_block(P.out_blocks.size());
  std::vector<int> tgroup;   // column strip of each stage-2 tile
  for (int q : order) {
    const Contrib& c = P.contrib[q];
    if (c.kind == C_BOTH && lead[q] >= 0 && lead[q] != q) continue;   // inside its group's product
    PSeg s{};
    s.c = (c.kind == C_BOTH && lead[q] >= 0) ? 1.0 : c.c;
    switch (c.kind) {
      case C_AXPY:
        s.l_off = c.x_off; s.l_stride = c.x_ld; s.spaces = PK_XF; s.K = 0;
        break;
      case C_LEFT: {
        const Line a = left_operand(ops, c.a_off, c.a_ld, c.a_t, c.M, c.K1);
        const Line x = right_operand(xs, c.x_off, c.x_ld, 0, c.K1, c.N);
        s.l_off = a.off; s.l_stride = a.stride;
        s.r_off = x.off; s.r_stride = x.stride;
        s.spaces = PK_OPS | (PK_X << 8);
        s.K = pad16(c.K1);
        break;
      }
      case C_RIGHT: {
        const Line x = left_operand(xs, c.x_off, c.x_ld, 0, c.M, c.K2);
        const Line b = right_operand(ops, c.b_off, c.b_ld, c.b_t, c.K2, c.N);
        s.l_off = x.off; s.l_stride = x.stride;
        s.r_off = b.off; s.r_stride = b.stride;
        s.spaces = PK_X | (PK_OPS << 8);
        s.K = pad16(c.K2);
        break;
      }
      default: {   // C_BOTH: T op(B)
        const TRef t = t_of[lead[q] >= 0 ? lead[q] : q];
        const Line b = right_operand(ops, c.b_off, c.b_ld, c.b_t, c.K2, c.N);
        s.l_off = t.off; s.l_stride = t.stride;
        s.r_off = b.off; s.r_stride = b.stride;
        s.spaces = PK_T | (PK_OPS << 8);
        s.K = pad16(c.K2);
        break;
      }
    }
    per_block[c.out_block].push_back(s);
  }
  for (size_t b = 0; b < P.out_blocks.size(); ++b) {
    const OutBlock& ob = P.out_blocks[b];
    const int Mp = pad16(ob.M), Np = pad16(ob.N);
    std::vector<PSeg> pieces;
    for (const PSeg& s : per_block[b])
      if (s.l_space() == PK_XF) pieces.push_back(s);
    for (const PSeg& s0 : per_block[b]) {
      if (s0.l_space() == PK_XF) continue;
      for (int k_lo = 0; k_lo < s0.K; k_lo += shape.k_split) {
        PSeg s = s0;
        s.K = std::min(shape.k_split, s0.K - k_lo);
        s.l_off += (int64_t)k_lo * 4;
        s.r_off += (int64_t)k_lo * 4;
        pieces.push_back(s);
      }
    }
    // work items of at most item_chunks K chunks (an axpy counts one)
    const int first = (int)out.seg2.size();
    out.seg2.insert(out.seg2.end(), pieces.begin(), pieces.end());
    std::vector<std::pair<int, int>> items;
    int ib = first, cost = 0;
    for (int q = first; q < (int)out.seg2.size(); ++q) {
      const int c = out.seg2[q].l_space() == PK_XF ? 1 : (out.seg2[q].K + 31) / 32;
      if (q > ib && cost + c > shape.item_chunks) {
        items.push_back({ib, q});
        ib = q;
        cost = 0;
      }
      cost += c;
    }
    items.push_back({ib, (int)out.seg2.size()});   // also the empty item of a block without terms
    for (int m0 = 0; m0 < Mp; m0 += 64)
      for (int n0 = 0; n0 < Np; n0 += 64) {
        PTile tl{};
        tl.out_off = ob.off; tl.out_ld = ob.ld;
        tl.m0 = m0; tl.n0 = n0;
        tl.mb = std::min(64, Mp - m0) / 16; tl.nb = std::min(64, Np - n0) / 16;
        tl.mlim = ob.M; tl.nlim = ob.N;
        tl.out_kind = PK_OUT_Y;
        if (items.size() == 1) {
          tl.seg_begin = items[0].first; tl.seg_end = items[0].second; tl.slot = -1;
          out.tile2.push_back(tl);
          tgroup.push_back(ob.group);
          continue;
        }
        const int s0 = (int)out.nslots;
        for (auto& it : items) {
          tl.seg_begin = it.first; tl.seg_end = it.second; tl.slot = (int)out.nslots++;
          out.tile2.push_back(tl);
          tgroup.push_back(ob.group);
        }
        PReduce r{};
        r.out_off = ob.off; r.out_ld = ob.ld; r.m0 = m0; r.n0 = n0;
        r.mt = std::min(64, ob.M - m0); r.nt = std::min(64, ob.N - n0);
        r.s0 = s0; r.s1 = (int)out.nslots;
        out.rtile.push_back(r);
      }
  }
  if (shape.xcds2 > 1 && !out.tile2.empty()) {
    // locality group: the left operand of the tile's first GEMM segment (a
    // T block, or an operator or x block), which the tiles of several
    // output blocks read; a tile of axpy segments only, its column strip
    std::map<int64_t, int> gid;
    std::vector<int> g2(out.tile2.size());
    for (size_t t = 0; t < out.tile2.size(); ++t) {
      int64_t k = -1 - (int64_t)tgroup[t];
      for (int q = out.tile2[t].seg_begin; q < out.tile2[t].seg_end; ++q)
        if (out.seg2[q].l_space() != PK_XF) {
          k = out.seg2[q].l_off * 4 + out.seg2[q].l_space();
          break;
        }
      g2[t] = gid.emplace(k, (int)gid.size()).first->second;
    }
    xcd_bins(out.tile2, g2, tile_costs(out.tile2, out.seg2), shape.xcds2, PK_OUT_Y);
  }
  if (ops.bad || xs.bad) {
    msg = "packed layout: one operand offset referenced with two shapes";
    return TJSQ_ERR_ARG;
  }
  // MFMA flops the tiles execute (K in chunks of 32, tiles in 16 x 16 blocks)
  double f1 = 0;
  for (const PTile& t : out.tile1)
    for (int q = t.seg_begin; q < t.seg_end; ++q) f1 += 2.0 * (16 * t.mb) * (16 * t.nb) * ((out.seg1[q].K + 31) / 32 * 32);
  double f2 = 0;
  for (const PTile& t : out.tile2)
    for (int q = t.seg_begin; q < t.seg_end; ++q)
      if (out.seg2[q].l_space() != PK_XF) f2 += 2.0 * (16 * t.mb) * (16 * t.nb) * ((out.seg2[q].K + 31) / 32 * 32);
  out.flops_exec = f1 + f2;
  return TJSQ_OK;
}

}  // namespace tjsq
