// libpgtree.so: host-side (CPU, OpenMP) construction of the interaction lists
// that the batched GPU calls consume.  This is caller-side code: the
// reference ships its Barnes-Hut list builder inside the sample program
// (PG5/newton_quad/source.cpp:215-354 moments, :581-707 opening test and list
// walk) and no TreePM list builder at all.
//
//  * pgt_treepm_build: periodic unit box; ghost copies of the particles within
//    r_cut of a face (so the GPU kernel needs no minimum-image logic);
//    i-groups = octree nodes of <= ncrit particles; j-list of a group = every
//    particle or ghost within r_cut of the group's bounding box, in the order
//    of a uniform grid of cells of side >= r_cut (cells by x, y, z index, a
//    cell's particles by index).
//  * pgt_bh_build: open domain; i-groups = octree nodes of <= ncrit particles;
//    the tree walk of source.cpp:660-707 with its opening test (a child is
//    accepted as a cell when theta^2 * d^2 > s^2, d = distance from the child's
//    centre of mass to the group box, s = child side), children holding one
//    particle and leaves (<= nleaf particles) contributing particles; cells
//    carry mass, centre of mass and the traceless quadrupole
//    sum m (3 d d - |d|^2 I) about the centre of mass.
//
// Results live in an opaque handle: pgt_info returns the sizes, pgt_get
// copies one array, pgt_free releases it.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#define PGT_API extern "C" __attribute__((visibility("default")))

namespace {

struct Result {
  // particle arrays (extended set for TreePM: originals then ghosts)
  std::vector<double> pos;   // [np][3]
  std::vector<int> src;      // original index of each entry of pos
  std::vector<int> ioff, icnt, joff, jcnt, coff, ccnt, order;
  std::vector<int> ilist, jlist, clist;
  std::vector<double> cells;  // [ncells][10]: x, y, z, M, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz
  long long work_pp = 0, work_pc = 0;  // sum ni*nj, sum ni*nc
};

uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

struct Node {
  int lo, hi;        // particle range in Morton order
  int child[8];      // -1 if empty
  double c[3], l;    // geometric centre and side
  double com[3], m;  // centre of mass and mass
  double q[6];       // traceless quadrupole about com
  int cell = -1;     // index in the cell array (internal nodes)
};

inline bool is_leaf(const Node& nd) {
  for (int o = 0; o < 8; ++o)
    if (nd.child[o] >= 0) return false;
  return true;
}

struct Octree {
  std::vector<int> perm;  // Morton order -> input index
  std::vector<uint64_t> key;
  std::vector<Node> nodes;
  std::vector<int> groups;  // node indices of the i-groups
};

// Octree over n points inside the cube [lo, lo + L)^3, split while a node
// holds more than `leaf` points (21 levels at most).
void build_octree(int n, const double* x, const double* m, const double lo[3], double L,
                  int ncrit, int nleaf, Octree& t) {
  t.perm.resize(n);
  t.key.resize(n);
  std::vector<std::pair<uint64_t, int>> kv(n);
  const double s = (double)(1 << 21) / L;
#pragma omp parallel for
  for (int i = 0; i < n; ++i) {
    uint64_t k = 0;
    for (int d = 0; d < 3; ++d) {
      double u = (x[3 * i + d] - lo[d]) * s;
      int64_t q = (int64_t)std::floor(u);
      q = std::min<int64_t>(std::max<int64_t>(q, 0), (1 << 21) - 1);
      k |= spread3((uint64_t)q) << (2 - d);
    }
    kv[i] = {k, i};
  }
  std::sort(kv.begin(), kv.end());
  for (int i = 0; i < n; ++i) {
    t.key[i] = kv[i].first;
    t.perm[i] = kv[i].second;
  }
  t.nodes.clear();
  t.groups.clear();
  // iterative construction
  struct Todo {
    int node, level;
  };
  Node root;
  root.lo = 0;
  root.hi = n;
  for (int d = 0; d < 3; ++d) root.c[d] = lo[d] + 0.5 * L;
  root.l = L;
  std::fill(root.child, root.child + 8, -1);
  t.nodes.push_back(root);
  std::vector<Todo> stack{{0, 0}};
  const int split = std::max(1, std::min(ncrit, nleaf));
  while (!stack.empty()) {
    Todo td = stack.back();
    stack.pop_back();
    Node nd = t.nodes[td.node];
    const int cnt = nd.hi - nd.lo;
    if (cnt <= split || td.level >= 21) continue;
    const int shift = 3 * (20 - td.level);
    int a = nd.lo;
    for (int oct = 0; oct < 8; ++oct) {
      int b = a;
      while (b < nd.hi && (int)((t.key[b] >> shift) & 7) == oct) ++b;
      if (b > a) {
        Node ch;
        ch.lo = a;
        ch.hi = b;
        ch.l = 0.5 * nd.l;
        // octant bits: x is bit 2, y bit 1, z bit 0 of each 3-bit digit
        ch.c[0] = nd.c[0] + ((oct & 4) ? 0.25 : -0.25) * nd.l;
        ch.c[1] = nd.c[1] + ((oct & 2) ? 0.25 : -0.25) * nd.l;
        ch.c[2] = nd.c[2] + ((oct & 1) ? 0.25 : -0.25) * nd.l;
        std::fill(ch.child, ch.child + 8, -1);
        const int id = (int)t.nodes.size();
        t.nodes.push_back(ch);
        t.nodes[td.node].child[oct] = id;
        stack.push_back({id, td.level + 1});
      }
      a = b;
    }
  }
  // Moments, bottom-up (children are created after their parent, so a reverse
  // sweep sees every child first): a leaf sums its particles; an internal
  // node combines its children's mass, centre of mass and quadrupole with the
  // parallel-axis shift.  The GPU builder (treebuild_dev.hip) evaluates the
  // same operations in the same order, so both give identical bits.
  const int nn = (int)t.nodes.size();
  for (int k = nn - 1; k >= 0; --k) {
    Node& nd = t.nodes[k];
    bool leaf = true;
    for (int o = 0; o < 8; ++o) leaf = leaf && nd.child[o] < 0;
    double M = 0, c[3] = {0, 0, 0}, q[6] = {0, 0, 0, 0, 0, 0};
    if (leaf) {
      for (int p = nd.lo; p < nd.hi; ++p) {
        const int i = t.perm[p];
        const double mi = m ? m[i] : 1.0;
        M += mi;
        for (int d = 0; d < 3; ++d) c[d] += mi * x[3 * i + d];
      }
      for (int d = 0; d < 3; ++d) nd.com[d] = M > 0 ? c[d] / M : nd.c[d];
      for (int p = nd.lo; p < nd.hi; ++p) {
        const int i = t.perm[p];
        const double mi = m ? m[i] : 1.0;
        double dd[3];
        for (int d = 0; d < 3; ++d) dd[d] = x[3 * i + d] - nd.com[d];
        const double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        int kk = 0;
        for (int a = 0; a < 2; ++a)
          for (int b = a; b < 3; ++b) {
            double v = 3.0 * dd[a] * dd[b];
            if (a == b) v -= r2;
            q[kk++] += mi * v;
          }
      }
    } else {
      for (int o = 0; o < 8; ++o) {
        if (nd.child[o] < 0) continue;
        const Node& ch = t.nodes[nd.child[o]];
        M += ch.m;
        for (int d = 0; d < 3; ++d) c[d] += ch.m * ch.com[d];
      }
      for (int d = 0; d < 3; ++d) nd.com[d] = M > 0 ? c[d] / M : nd.c[d];
      for (int o = 0; o < 8; ++o) {
        if (nd.child[o] < 0) continue;
        const Node& ch = t.nodes[nd.child[o]];
        double dd[3];
        for (int d = 0; d < 3; ++d) dd[d] = ch.com[d] - nd.com[d];
        const double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        int kk = 0;
        for (int a = 0; a < 2; ++a)
          for (int b = a; b < 3; ++b) {
            double v = 3.0 * dd[a] * dd[b];
            if (a == b) v -= r2;
            q[kk] += ch.q[kk] + ch.m * v;
            ++kk;
          }
      }
    }
    nd.m = M;
    q[5] = -(q[0] + q[3]);
    std::memcpy(nd.q, q, sizeof(q));
  }
  // i-groups: nodes with <= ncrit particles whose parent has more, and nodes
  // the 21-level key cannot split (more than ncrit particles inside one cell
  // of side L / 2^21), which form one group of all their particles
  std::vector<int> st{0};
  while (!st.empty()) {
    const int k = st.back();
    st.pop_back();
    const Node& nd = t.nodes[k];
    if (nd.hi - nd.lo <= ncrit || is_leaf(nd)) {
      t.groups.push_back(k);
      continue;
    }
    for (int o = 7; o >= 0; --o)
      if (nd.child[o] >= 0) st.push_back(nd.child[o]);
  }
  std::sort(t.groups.begin(), t.groups.end(),
            [&](int a, int b) { return t.nodes[a].lo < t.nodes[b].lo; });
}

void finish_order(Result& r, bool cells) {
  const int ng = (int)r.ioff.size();
  std::vector<std::pair<long long, int>> c(ng);
  for (int g = 0; g < ng; ++g)
    c[g] = {-(long long)r.icnt[g] * (r.jcnt[g] + (cells ? 2 * r.ccnt[g] : 0)), g};
  std::stable_sort(c.begin(), c.end());
  r.order.resize(ng);
  for (int g = 0; g < ng; ++g) r.order[g] = c[g].second;
  r.work_pp = r.work_pc = 0;
  for (int g = 0; g < ng; ++g) {
    r.work_pp += (long long)r.icnt[g] * r.jcnt[g];
    if (cells) r.work_pc += (long long)r.icnt[g] * r.ccnt[g];
  }
}

// Concatenate per-group lists into CSR-like (offset, count) arrays.
void concat(const std::vector<std::vector<int>>& per, std::vector<int>& off,
            std::vector<int>& cnt, std::vector<int>& flat) {
  const int ng = (int)per.size();
  off.resize(ng);
  cnt.resize(ng);
  size_t tot = 0;
  for (int g = 0; g < ng; ++g) {
    off[g] = (int)tot;
    cnt[g] = (int)per[g].size();
    tot += per[g].size();
  }
  flat.resize(tot);
#pragma omp parallel for
  for (int g = 0; g < ng; ++g) std::copy(per[g].begin(), per[g].end(), flat.begin() + off[g]);
}

}  // namespace

// TreePM lists for n particles in the periodic unit box (positions in [0,1)).
PGT_API void* pgt_treepm_build(int n, const double* x, double rcut, int ncrit) {
  Result* r = new Result();
  // originals in Morton order of the octree, then ghosts
  Octree t;
  const double lo[3] = {0, 0, 0};
  build_octree(n, x, nullptr, lo, 1.0, ncrit, ncrit, t);
  for (int p = 0; p < n; ++p) {
    const int i = t.perm[p];
    r->pos.insert(r->pos.end(), {x[3 * i], x[3 * i + 1], x[3 * i + 2]});
    r->src.push_back(i);
  }
  r->pos.reserve(3 * (size_t)n * 8);  // at most 7 ghosts per particle
  r->src.reserve((size_t)n * 8);
  for (int p = 0; p < n; ++p) {
    // a copy: appending to r->pos below may move its storage
    const double xp[3] = {r->pos[3 * p], r->pos[3 * p + 1], r->pos[3 * p + 2]};
    int lo_s[3], hi_s[3];
    for (int d = 0; d < 3; ++d) {
      lo_s[d] = xp[d] >= 1.0 - rcut ? -1 : 0;
      hi_s[d] = xp[d] < rcut ? 1 : 0;
    }
    for (int sx = lo_s[0]; sx <= hi_s[0]; ++sx)
      for (int sy = lo_s[1]; sy <= hi_s[1]; ++sy)
        for (int sz = lo_s[2]; sz <= hi_s[2]; ++sz) {
          if (!sx && !sy && !sz) continue;
          r->pos.insert(r->pos.end(), {xp[0] + sx, xp[1] + sy, xp[2] + sz});
          r->src.push_back(r->src[p]);
        }
  }
  const int np = (int)r->src.size();
  // uniform grid over [-rcut, 1 + rcut) with cells of side >= rcut
  const double glo = -rcut, gL = 1.0 + 2 * rcut;
  const int nc = std::max(1, (int)std::floor(gL / rcut));
  const double h = gL / nc;
  auto cidx = [&](double v) {
    int c = (int)std::floor((v - glo) / h);
    return std::min(std::max(c, 0), nc - 1);
  };
  std::vector<int> cstart((size_t)nc * nc * nc + 1, 0), cell_of(np);
  for (int p = 0; p < np; ++p) {
    const double* xp = &r->pos[3 * p];
    cell_of[p] = (cidx(xp[0]) * nc + cidx(xp[1])) * nc + cidx(xp[2]);
    cstart[cell_of[p] + 1]++;
  }
  for (size_t c = 0; c + 1 < cstart.size(); ++c) cstart[c + 1] += cstart[c];
  std::vector<int> cpart(np), fill(cstart.begin(), cstart.end() - 1);
  for (int p = 0; p < np; ++p) cpart[fill[cell_of[p]]++] = p;
  // groups (group boxes = bounding boxes of their particles)
  const int ng = (int)t.groups.size();
  std::vector<std::vector<int>> il(ng), jl(ng);
  const double rc2 = rcut * rcut;
#pragma omp parallel for schedule(dynamic, 8)
  for (int g = 0; g < ng; ++g) {
    const Node& nd = t.nodes[t.groups[g]];
    double bmin[3] = {1e30, 1e30, 1e30}, bmax[3] = {-1e30, -1e30, -1e30};
    for (int p = nd.lo; p < nd.hi; ++p) {
      il[g].push_back(p);
      for (int d = 0; d < 3; ++d) {
        bmin[d] = std::min(bmin[d], r->pos[3 * p + d]);
        bmax[d] = std::max(bmax[d], r->pos[3 * p + d]);
      }
    }
    int c0[3], c1[3];
    for (int d = 0; d < 3; ++d) {
      c0[d] = cidx(bmin[d] - rcut);
      c1[d] = cidx(bmax[d] + rcut);
    }
    for (int a = c0[0]; a <= c1[0]; ++a)
      for (int b = c0[1]; b <= c1[1]; ++b)
        for (int c = c0[2]; c <= c1[2]; ++c) {
          const int cc = (a * nc + b) * nc + c;
          for (int k = cstart[cc]; k < cstart[cc + 1]; ++k) {
            const int p = cpart[k];
            double d2 = 0;
            for (int d = 0; d < 3; ++d) {
              const double v = r->pos[3 * p + d];
              const double e = v < bmin[d] ? bmin[d] - v : (v > bmax[d] ? v - bmax[d] : 0.0);
              d2 += e * e;
            }
            if (d2 < rc2) jl[g].push_back(p);
          }
        }
  }
  concat(il, r->ioff, r->icnt, r->ilist);
  concat(jl, r->joff, r->jcnt, r->jlist);
  r->coff.assign(ng, 0);
  r->ccnt.assign(ng, 0);
  finish_order(*r, false);
  return r;
}

// Barnes-Hut lists for n particles (open domain), masses m.
PGT_API void* pgt_bh_build(int n, const double* x, const double* m, double theta, int ncrit,
                           int nleaf) {
  Result* r = new Result();
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], x[3 * i + d]);
      hi[d] = std::max(hi[d], x[3 * i + d]);
    }
  double L = 0;
  for (int d = 0; d < 3; ++d) L = std::max(L, hi[d] - lo[d]);
  L *= 1.0 + 1e-9;
  Octree t;
  build_octree(n, x, m, lo, L, ncrit, nleaf, t);
  for (int p = 0; p < n; ++p) {
    const int i = t.perm[p];
    r->pos.insert(r->pos.end(), {x[3 * i], x[3 * i + 1], x[3 * i + 2]});
    r->src.push_back(i);
  }
  // cells: every internal node with more than one particle
  for (size_t k = 0; k < t.nodes.size(); ++k) {
    Node& nd = t.nodes[k];
    if (nd.hi - nd.lo <= 1) continue;
    nd.cell = (int)(r->cells.size() / 10);
    r->cells.insert(r->cells.end(), {nd.com[0], nd.com[1], nd.com[2], nd.m, nd.q[0], nd.q[1],
                                     nd.q[2], nd.q[3], nd.q[4], nd.q[5]});
  }
  const double th2 = theta * theta;
  const int ng = (int)t.groups.size();
  std::vector<std::vector<int>> il(ng), pl(ng), cl(ng);
#pragma omp parallel for schedule(dynamic, 4)
  for (int g = 0; g < ng; ++g) {
    const Node& gn = t.nodes[t.groups[g]];
    for (int p = gn.lo; p < gn.hi; ++p) il[g].push_back(p);
    const double li = 0.5 * gn.l;
    auto box_d2 = [&](const double* c) {
      double d2 = 0;
      for (int d = 0; d < 3; ++d) {
        const double e = std::max(std::fabs(c[d] - gn.c[d]) - li, 0.0);
        d2 += e * e;
      }
      return d2;
    };
    const Node& root = t.nodes[0];
    if (box_d2(root.com) * th2 >= root.l * root.l) {  // source.cpp:832-840
      cl[g].push_back(root.cell);
      continue;
    }
    std::vector<int> st{0};
    while (!st.empty()) {
      const Node& nd = t.nodes[st.back()];
      st.pop_back();
      if (nd.hi - nd.lo <= nleaf || is_leaf(nd)) {  // a node the key cannot split lists its particles
        for (int p = nd.lo; p < nd.hi; ++p) pl[g].push_back(p);
        continue;
      }
      const double s2 = 0.25 * nd.l * nd.l;  // child side squared
      for (int o = 7; o >= 0; --o) {
        const int ck = nd.child[o];
        if (ck < 0) continue;
        const Node& ch = t.nodes[ck];
        if (ch.hi - ch.lo == 1) {
          pl[g].push_back(ch.lo);
        } else if (box_d2(ch.com) * th2 > s2) {
          cl[g].push_back(ch.cell);
        } else {
          st.push_back(ck);
        }
      }
    }
  }
  concat(il, r->ioff, r->icnt, r->ilist);
  concat(pl, r->joff, r->jcnt, r->jlist);
  concat(cl, r->coff, r->ccnt, r->clist);
  finish_order(*r, true);
  return r;
}

// info[10] = {np, ngroups, len(ilist), len(jlist), len(clist), ncells,
//             work_pp (sum ni*nj) as two ints lo/hi, work_pc lo/hi}
PGT_API int pgt_info(void* h, long long* info) {
  const Result* r = static_cast<Result*>(h);
  info[0] = (long long)r->src.size();
  info[1] = (long long)r->ioff.size();
  info[2] = (long long)r->ilist.size();
  info[3] = (long long)r->jlist.size();
  info[4] = (long long)r->clist.size();
  info[5] = (long long)(r->cells.size() / 10);
  info[6] = r->work_pp;
  info[7] = r->work_pc;
  return 0;
}

// what: 0 pos (double [np][3]), 1 src, 2 ioff, 3 icnt, 4 ilist, 5 joff,
// 6 jcnt, 7 jlist, 8 coff, 9 ccnt, 10 clist, 11 order, 12 cells (double [nc][10])
PGT_API int pgt_get(void* h, int what, void* out) {
  const Result* r = static_cast<Result*>(h);
  auto cp = [&](const auto& v) {
    std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
    return 0;
  };
  switch (what) {
    case 0: return cp(r->pos);
    case 1: return cp(r->src);
    case 2: return cp(r->ioff);
    case 3: return cp(r->icnt);
    case 4: return cp(r->ilist);
    case 5: return cp(r->joff);
    case 6: return cp(r->jcnt);
    case 7: return cp(r->jlist);
    case 8: return cp(r->coff);
    case 9: return cp(r->ccnt);
    case 10: return cp(r->clist);
    case 11: return cp(r->order);
    case 12: return cp(r->cells);
    default: return -1;
  }
}

PGT_API void pgt_free(void* h) { delete static_cast<Result*>(h); }
