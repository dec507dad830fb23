// semx_setup.cpp -- host-side setup of the spectral-element Helmholtz core:
// GLL operators, element geometry, global numbering, elemental static
// condensation and banded Cholesky factorisation.  This is the work of
// semtex's Element / Mesh / BCmgr / MatrixSys constructors; it runs once
// per system and is excluded from solve timings.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>
#include <utility>

#include "semx_internal.h"

namespace semx {

// ---------------------------------------------------------------------------
// Quadrature: Gauss-Lobatto-Legendre points/weights and the derivative
// matrix of the Lagrange interpolants (femlib/polylib.c:312-340, 513-556).
// ---------------------------------------------------------------------------

static void legendre_pd(int n, long double x, long double& p, long double& d) {
  long double pm = 1.0L, pc = x, dm = 0.0L, dc = 1.0L;
  if (n == 0) { p = 1.0L; d = 0.0L; return; }
  for (int k = 2; k <= n; ++k) {
    long double pn = ((2 * k - 1) * x * pc - (k - 1) * pm) / k;
    long double dn = dm + (2 * k - 1) * pc;
    pm = pc; pc = pn; dm = dc; dc = dn;
  }
  p = pc; d = dc;
}

void gll(int np, std::vector<double>& z, std::vector<double>& w) {
  const int N = np - 1;
  z.assign(np, 0.0); w.assign(np, 0.0);
  if (np == 1) { w[0] = 2.0; return; }
  z[0] = -1.0; z[N] = 1.0;
  // Interior nodes are the roots of L'_N; Newton from Chebyshev-Lobatto
  // guesses, negative half only, mirrored for exact antisymmetry.
  for (int j = 1; 2 * j <= N; ++j) {
    long double x = -std::cos(3.14159265358979323846264338327950288L * j / N);
    for (int it = 0; it < 100; ++it) {
      long double p, d;
      legendre_pd(N, x, p, d);
      long double d2 = (2.0L * x * d - N * (N + 1.0L) * p) / (1.0L - x * x);
      long double dx = d / d2;
      x -= dx;
      if (std::fabs(dx) < 1e-19L) break;
    }
    z[j] = static_cast<double>(x);
    z[N - j] = -z[j];
  }
  if (N % 2 == 0) z[N / 2] = 0.0;
  for (int j = 0; j <= N; ++j) {
    long double p, d;
    legendre_pd(N, z[j], p, d);
    w[j] = static_cast<double>(2.0L / (N * (N + 1.0L) * p * p));
  }
}

void gll_deriv(int np, const std::vector<double>& z, std::vector<double>& D) {
  const int N = np - 1;
  D.assign(np * np, 0.0);
  std::vector<long double> L(np);
  for (int i = 0; i < np; ++i) { long double d; legendre_pd(N, z[i], L[i], d); }
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      if (i != j)
        D[i * np + j] = static_cast<double>(L[i] / (L[j] * ((long double)z[i] - z[j])));
    }
  if (np > 1) {
    D[0] = -0.25 * N * (N + 1);
    D[N * np + N] = 0.25 * N * (N + 1);
  }
}

// SVV-stabilised derivative SDV = BW diag(sqrt S) FW DV with the Legendre
// transform pair of femlib/operators.c:317-480 (src/svv.cpp:126-185).
void svv_derivative(int np, int svv_mn, double svv_eps, double kinvis,
                    std::vector<double>& SDV) {
  const int N = np - 1;
  std::vector<double> z, w, DV;
  gll(np, z, w);
  gll_deriv(np, z, DV);
  std::vector<double> Lt(np * np), S(np, 1.0);  // Lt[k*np+j] = L_k(z_j)
  for (int k = 0; k < np; ++k)
    for (int j = 0; j < np; ++j) {
      long double p, d;
      legendre_pd(k, z[j], p, d);
      Lt[k * np + j] = static_cast<double>(p);
    }
  if (svv_mn >= 0 && svv_mn < N && svv_eps > EPSDP) {
    for (int i = 0; i < np; ++i) S[i] = 0.0;
    for (int i = svv_mn + 1; i < np; ++i) {
      const double q = (N - i) / (1.0 * svv_mn - i);
      S[i] = std::exp(-q * q);
    }
    for (int i = 0; i < np; ++i) S[i] = 1.0 + svv_eps / kinvis * S[i];
  }
  // A = diag(sqrt S) FW, FW[k][j] = C_k w_j L_k(z_j)
  std::vector<double> A(np * np), B(np * np);
  for (int k = 0; k < np; ++k) {
    const double ck = (k < N) ? 0.5 * (2 * k + 1) : 0.5 * N;
    for (int j = 0; j < np; ++j) A[k * np + j] = std::sqrt(S[k]) * ck * w[j] * Lt[k * np + j];
  }
  // B = BW A, BW[i][k] = L_k(z_i)
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      double s = 0.0;
      for (int k = 0; k < np; ++k) s += Lt[k * np + i] * A[k * np + j];
      B[i * np + j] = s;
    }
  SDV.assign(np * np, 0.0);
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      double s = 0.0;
      for (int k = 0; k < np; ++k) s += B[i * np + k] * DV[k * np + j];
      SDV[i * np + j] = s;
    }
}

static void build_ops(Ops& o, int np, int svv_mn, double svv_eps, double kinvis) {
  o.np = np;
  gll(np, o.z, o.w);
  gll_deriv(np, o.z, o.DV);
  o.DT.assign(np * np, 0.0);
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) o.DT[j * np + i] = o.DV[i * np + j];
  svv_derivative(np, svv_mn, svv_eps, kinvis, o.SDV);
  o.SDT.assign(np * np, 0.0);
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) o.SDT[j * np + i] = o.SDV[i * np + j];
  // Edge-major permutation (femlib/mapping.c:46-132): CCW boundary from
  // node 0, then interior row-major.
  o.emap.clear();
  const int nm = np - 1;
  for (int j = 0; j < nm; ++j) o.emap.push_back(j);                 // bottom
  for (int i = 0; i < nm; ++i) o.emap.push_back(i * np + nm);       // right
  for (int j = nm; j > 0; --j) o.emap.push_back(nm * np + j);       // top
  for (int i = nm; i > 0; --i) o.emap.push_back(i * np);            // left
  for (int i = 1; i < nm; ++i)
    for (int j = 1; j < nm; ++j) o.emap.push_back(i * np + j);
  o.pmap.assign(np * np, 0);
  for (int k = 0; k < np * np; ++k) o.pmap[o.emap[k]] = k;
}

// ---------------------------------------------------------------------------
// Element geometry (src/mesh.cpp:772-886, src/element.cpp:1281-1436).
// ---------------------------------------------------------------------------

static void element_nodes(const Ops& o, const double* vx, const double* vy,
                          double* x, double* y) {
  const int np = o.np;
  // Straight sides: the Coons patch of straight sides is the bilinear map
  // of the four vertices evaluated at the GLL node pairs.
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      const double r = o.z[j], s = o.z[i];
      const double n0 = 0.25 * (1 - r) * (1 - s), n1 = 0.25 * (1 + r) * (1 - s);
      const double n2 = 0.25 * (1 + r) * (1 + s), n3 = 0.25 * (1 - r) * (1 + s);
      x[i * np + j] = n0 * vx[0] + n1 * vx[1] + n2 * vx[2] + n3 * vx[3];
      y[i * np + j] = n0 * vy[0] + n1 * vy[1] + n2 * vy[2] + n3 * vy[3];
    }
}

// Geometric factors of a straight-sided quad.  x(r,s) is the bilinear map
// ax + bx r + cx s + dx r s of the vertices, so its partials are evaluated in
// closed form (the quantities src/element.cpp:1281-1436 forms with the GLL
// derivative matrix); axis-aligned rectangles then give exactly zero cross
// terms.
static int element_factors(const Ops& o, const double* vx, const double* vy, double* g,
                           const double* ynode = nullptr) {
  const int np = o.np, npnp = np * np;
  const double bx = 0.25 * ((vx[1] - vx[0]) + (vx[2] - vx[3]));
  const double cx = 0.25 * ((vx[3] - vx[0]) + (vx[2] - vx[1]));
  const double dx = 0.25 * ((vx[0] - vx[1]) + (vx[2] - vx[3]));
  const double by = 0.25 * ((vy[1] - vy[0]) + (vy[2] - vy[3]));
  const double cy = 0.25 * ((vy[3] - vy[0]) + (vy[2] - vy[1]));
  const double dy = 0.25 * ((vy[0] - vy[1]) + (vy[2] - vy[3]));
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      const double r = o.z[j], s = o.z[i];
      const double xr = bx + dx * s, xs = cx + dx * r;
      const double yr = by + dy * s, ys = cy + dy * r;
      const double jac = xr * ys - xs * yr;
      if (jac < 4 * EPSDP) return SEMX_E_JACOBIAN;
      const double W = o.w[i] * o.w[j];
      const int q = i * np + j;
      g[G_Q1 * npnp + q] = (xs * xs + ys * ys) / jac * W;
      g[G_Q2 * npnp + q] = (xr * xr + yr * yr) / jac * W;
      g[G_Q3 * npnp + q] = -(yr * ys + xr * xs) / jac * W;
      g[G_Q4 * npnp + q] = jac * W;
      g[G_DRDX * npnp + q] = ys / jac;
      g[G_DRDY * npnp + q] = -xs / jac;
      g[G_DSDX * npnp + q] = -yr / jac;
      g[G_DSDY * npnp + q] = xr / jac;
      g[G_Y * npnp + q] = 0.0;
      // cylindrical: the stiffness factors premultiplied by the radius, the
      // mass kept unweighted (src/element.cpp:1417-1426, _Q8)
      if (ynode) {
        g[G_Q1 * npnp + q] *= ynode[q];
        g[G_Q2 * npnp + q] *= ynode[q];
        g[G_Q3 * npnp + q] *= ynode[q];
        g[G_Y * npnp + q] = ynode[q];
      }
    }
  // Null-mapping: factors that vanish to roundoff are exact zeros
  // (src/element.cpp:1431-1435).
  const int slots[5] = {G_DRDX, G_DRDY, G_DSDX, G_DSDY, G_Q3};
  for (int s : slots) {
    double n2 = 0;
    for (int q = 0; q < npnp; ++q) n2 += g[s * npnp + q] * g[s * npnp + q];
    if (std::sqrt(n2) < 4 * EPSDP) std::fill(g + s * npnp, g + (s + 1) * npnp, 0.0);
  }
  return SEMX_OK;
}

// ---------------------------------------------------------------------------
// Numbering: connectivity of element sides, essential nodes last, RCM.
// ---------------------------------------------------------------------------

static int band_of(const std::vector<int>& btog, const std::vector<int>& bmask,
                   int nel, int next) {
  int nb = 0;
  for (int e = 0; e < nel; ++e) {
    int lo = INT32_MAX, hi = -1;
    for (int i = 0; i < next; ++i)
      if (!bmask[e * next + i]) {
        lo = std::min(lo, btog[e * next + i]);
        hi = std::max(hi, btog[e * next + i]);
      }
    if (hi >= 0) nb = std::max(nb, hi - lo);
  }
  return nb + 1;
}

struct Graph {
  std::vector<int> ptr, adj;
  int n = 0;
  int deg(int v) const { return ptr[v + 1] - ptr[v]; }
};

static std::vector<int> bfs_levels(const Graph& G, int root, std::vector<int>& level) {
  std::vector<int> order;
  order.reserve(G.n);
  level.assign(G.n, -1);
  level[root] = 0;
  order.push_back(root);
  for (size_t h = 0; h < order.size(); ++h) {
    const int v = order[h];
    for (int k = G.ptr[v]; k < G.ptr[v + 1]; ++k) {
      const int u = G.adj[k];
      if (level[u] < 0) { level[u] = level[v] + 1; order.push_back(u); }
    }
  }
  return order;
}

// George-Liu pseudo-peripheral node search followed by Cuthill-McKee from
// that root, per connected component; returned as the reversed order.
static std::vector<int> rcm(const Graph& G, int start) {
  std::vector<int> perm;
  perm.reserve(G.n);
  std::vector<char> done(G.n, 0);
  std::vector<int> level;
  int seed = start;
  while ((int)perm.size() < G.n) {
    while (seed < G.n && done[seed]) ++seed;
    if (seed >= G.n) seed = 0;
    while (done[seed]) ++seed;
    int root = seed;
    // pseudo-peripheral refinement (restricted to this component)
    int ecc = -1;
    for (int it = 0; it < 8; ++it) {
      std::vector<int> ord = bfs_levels(G, root, level);
      const int e = level[ord.back()];
      if (e <= ecc) break;
      ecc = e;
      int best = ord.back();
      for (int v : ord)
        if (level[v] == e && G.deg(v) < G.deg(best)) best = v;
      root = best;
    }
    // Cuthill-McKee from root
    std::vector<int> comp;
    comp.push_back(root);
    done[root] = 1;
    for (size_t h = 0; h < comp.size(); ++h) {
      const int v = comp[h];
      std::vector<int> nb;
      for (int k = G.ptr[v]; k < G.ptr[v + 1]; ++k)
        if (!done[G.adj[k]]) { done[G.adj[k]] = 1; nb.push_back(G.adj[k]); }
      std::stable_sort(nb.begin(), nb.end(),
                       [&](int a, int b) { return G.deg(a) < G.deg(b); });
      comp.insert(comp.end(), nb.begin(), nb.end());
    }
    perm.insert(perm.end(), comp.begin(), comp.end());
  }
  std::reverse(perm.begin(), perm.end());
  return perm;  // perm[newid] = oldid
}

static int find_root(std::vector<int>& par, int a) {
  while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; }
  return a;
}

// A numbering supplied by the caller (semtex's session.num, read by
// BCmgr::buildnum, src/bcmgr.cpp:586-824): it must join exactly the slots
// the connectivity joins, mask exactly the essential nodes and number the
// nsolve unknowns before the essential nodes.
struct GivenNumbering {
  const int* btog;
  const int* bmask;
  int nglobal, nsolve;
};

// ---------------------------------------------------------------------------
// The numbering of semtex's `enumerate` utility (utility/enumerate.cpp:
// 540-943), which writes session.num: the traverse numbering of
// Mesh::buildMap (src/mesh.cpp:889-1015; the provisional ids below), the
// unknowns moved before the essential nodes in their traverse order
// (Nsys::sortGid), then for levels 1-3 a reverse Cuthill-McKee order of the
// unknowns (George & Liu 1981, SPARSPAK ROOTLS/FNROOT/DEGREE/RCM as in
// femlib/netlib.f) over the adjacency lists of Nsys::connectivSC:
//   1: FNROOT from the first unknown, then RCM from the root it finds;
//   2: FNROOT from every tenth unknown, RCM from the root of least
//      global bandwidth;
//   3: RCM from every unknown, the root of least global bandwidth
//      (enumerate's default, and the regression scripts').
// The first root of least bandwidth wins.  Levels 2 and 3 cost one RCM per
// trial root.  Returns false when the unknowns' graph is disconnected (RCM
// from one root then numbers only its component).
namespace {

struct SparGraph {
  std::vector<int> xadj, adj;   // 0-based CSR, neighbours in insertion order
};

// ROOTLS: the level structure rooted at root over nodes with mask != 0;
// xls[k] is the start of level k in ls, xls[nlvl] the end.
void sp_rootls(const SparGraph& G, int root, std::vector<char>& mask, int& nlvl,
               std::vector<int>& xls, std::vector<int>& ls) {
  mask[root] = 0;
  ls[0] = root;
  nlvl = 0;
  int lvlend = 0, ccsize = 1;
  for (;;) {
    const int lbegin = lvlend;
    lvlend = ccsize;
    xls[nlvl++] = lbegin;
    for (int i = lbegin; i < lvlend; ++i)
      for (int j = G.xadj[ls[i]]; j < G.xadj[ls[i] + 1]; ++j) {
        const int v = G.adj[j];
        if (!mask[v]) continue;
        mask[v] = 0;
        ls[ccsize++] = v;
      }
    if (ccsize == lvlend) break;
  }
  xls[nlvl] = lvlend;
  for (int i = 0; i < ccsize; ++i) mask[ls[i]] = 1;
}

// FNROOT: a pseudo-peripheral node of root's component.
void sp_fnroot(const SparGraph& G, int& root, std::vector<char>& mask, std::vector<int>& xls,
               std::vector<int>& ls) {
  int nlvl = 0, nunlvl = 0;
  sp_rootls(G, root, mask, nlvl, xls, ls);
  const int ccsize = xls[nlvl];
  if (nlvl == 1 || nlvl == ccsize) return;
  for (;;) {
    const int jstrt = xls[nlvl - 1];
    int mindeg = ccsize;
    root = ls[jstrt];
    if (ccsize != jstrt + 1)
      for (int j = jstrt; j < ccsize; ++j) {
        const int v = ls[j];
        int ndeg = 0;
        for (int k = G.xadj[v]; k < G.xadj[v + 1]; ++k) ndeg += mask[G.adj[k]] != 0;
        if (ndeg < mindeg) { root = v; mindeg = ndeg; }
      }
    sp_rootls(G, root, mask, nunlvl, xls, ls);
    if (nunlvl <= nlvl) return;
    nlvl = nunlvl;
    if (nlvl >= ccsize) return;
  }
}

// RCM from root over the masked component: breadth-first, each node's new
// neighbours in order of increasing degree (insertion, ties keep adjacency
// order), the whole order reversed.  Returns the component size.
int sp_rcm(const SparGraph& G, int root, std::vector<char>& mask, std::vector<int>& perm,
           std::vector<int>& deg) {
  // DEGREE: degree within the masked graph, for root's component
  {
    std::vector<int>& ls = perm;
    std::vector<char> seen(mask.size(), 0);
    int n = 1;
    ls[0] = root;
    seen[root] = 1;
    for (int h = 0; h < n; ++h) {
      const int v = ls[h];
      int d = 0;
      for (int j = G.xadj[v]; j < G.xadj[v + 1]; ++j) {
        const int w = G.adj[j];
        if (!mask[w]) continue;
        ++d;
        if (!seen[w]) { seen[w] = 1; ls[n++] = w; }
      }
      deg[v] = d;
    }
  }
  mask[root] = 0;
  perm[0] = root;
  int lnbr = 1, lvlend = 0;
  do {
    const int lbegin = lvlend;
    lvlend = lnbr;
    for (int i = lbegin; i < lvlend; ++i) {
      const int v = perm[i];
      const int fnbr = lnbr;
      for (int j = G.xadj[v]; j < G.xadj[v + 1]; ++j) {
        const int w = G.adj[j];
        if (!mask[w]) continue;
        mask[w] = 0;
        perm[lnbr++] = w;
      }
      for (int k = fnbr + 1; k < lnbr; ++k) {
        const int w = perm[k];
        int l = k - 1;
        while (l >= fnbr && deg[perm[l]] > deg[w]) { perm[l + 1] = perm[l]; --l; }
        perm[l + 1] = w;
      }
    }
  } while (lnbr > lvlend);
  std::reverse(perm.begin(), perm.begin() + lnbr);
  return lnbr;
}

}  // namespace

static bool enumerate_numbering(const std::vector<int>& prov, const std::vector<int>& bmask,
                                int ng, int nunk, int nel, int next, int level,
                                std::vector<int>& btog) {
  // sortGid: unknowns first, each class in traverse order
  std::vector<char> ess(ng, 0);
  for (size_t i = 0; i < prov.size(); ++i) ess[prov[i]] = (char)bmask[i];
  std::vector<int> newid(ng);
  for (int v = 0, ku = 0, ke = nunk; v < ng; ++v) newid[v] = ess[v] ? ke++ : ku++;
  btog.resize(prov.size());
  for (size_t i = 0; i < prov.size(); ++i) btog[i] = newid[prov[i]];
  if (level < 1 || nunk == 0) return true;
  // connectivSC adjacency: element by element, slot by slot, first
  // appearance only
  std::vector<std::vector<int>> lst(nunk);
  for (int e = 0; e < nel; ++e) {
    const int* bt = &btog[(size_t)e * next];
    const int* mk = &bmask[(size_t)e * next];
    for (int i = 0; i < next; ++i) {
      if (mk[i]) continue;
      std::vector<int>& L = lst[bt[i]];
      for (int j = 0; j < next; ++j)
        if (j != i && !mk[j] && std::find(L.begin(), L.end(), bt[j]) == L.end())
          L.push_back(bt[j]);
    }
  }
  SparGraph G;
  G.xadj.assign(nunk + 1, 0);
  for (int v = 0; v < nunk; ++v) G.xadj[v + 1] = G.xadj[v] + (int)lst[v].size();
  G.adj.reserve(G.xadj[nunk]);
  for (int v = 0; v < nunk; ++v) G.adj.insert(G.adj.end(), lst[v].begin(), lst[v].end());
  std::vector<char> mask(nunk, 1);
  std::vector<int> perm(nunk), xls(nunk + 1), deg(nunk), inv(nunk);
  auto bandwidth = [&]() {   // Nsys::globalBandwidth of the order in perm
    for (int k = 0; k < nunk; ++k) inv[perm[k]] = k;
    int bw = 0;
    for (int e = 0; e < nel; ++e) {
      int lo = INT32_MAX, hi = -1;
      for (int i = 0; i < next; ++i) {
        if (bmask[(size_t)e * next + i]) continue;
        const int g = inv[btog[(size_t)e * next + i]];
        lo = std::min(lo, g);
        hi = std::max(hi, g);
      }
      if (hi >= 0) bw = std::max(bw, hi - lo);
    }
    return bw + 1;
  };
  int root = 0;
  if (level == 1) {
    sp_fnroot(G, root, mask, xls, perm);
  } else {
    int bwmin = INT32_MAX;
    for (int r = 0; r < nunk; r += (level == 2) ? 10 : 1) {
      int rt = r;
      std::fill(mask.begin(), mask.end(), 1);
      if (level == 2) sp_fnroot(G, rt, mask, xls, perm);
      if (sp_rcm(G, rt, mask, perm, deg) != nunk) return false;
      const int bw = bandwidth();
      if (bw < bwmin) { bwmin = bw; root = rt; }
    }
    std::fill(mask.begin(), mask.end(), 1);
  }
  if (sp_rcm(G, root, mask, perm, deg) != nunk) return false;
  for (int k = 0; k < nunk; ++k) inv[perm[k]] = k;
  for (size_t i = 0; i < btog.size(); ++i)
    if (btog[i] < nunk) btog[i] = inv[btog[i]];
  return true;
}

static int number_mesh(Mesh& M, const int* nbr, const int* skind, const GivenNumbering* given,
                       int enum_level) {
  const int np = M.np, nel = M.nel, next = M.next, nm = np - 1, ni = np - 2;
  const int nside = 4 * nel;
  // -- validate the side connectivity: joined sides pair up symmetrically,
  //    every unjoined side carries a boundary kind
  for (int i = 0; i < nside; ++i) {
    const int j = nbr[i];
    if (j >= 0) {
      if (j >= nside || j == i || nbr[j] != i || skind[i] != 0) return SEMX_E_MESH;
    } else if (skind[i] != 1 && skind[i] != 2) {
      return SEMX_E_MESH;
    }
  }
  // -- corners of joined sides coincide: side (e,s) runs from corner s to
  //    corner s+1, its partner runs the other way (src/mesh.cpp:889-972)
  std::vector<int> par(nside);
  std::iota(par.begin(), par.end(), 0);
  for (int i = 0; i < nside; ++i) {
    const int j = nbr[i];
    if (j <= i) continue;
    const int e = i / 4, s = i % 4, e2 = j / 4, s2 = j % 4;
    const int a0 = find_root(par, 4 * e + s), b0 = find_root(par, 4 * e2 + (s2 + 1) % 4);
    par[a0] = b0;
    const int a1 = find_root(par, 4 * e + (s + 1) % 4), b1 = find_root(par, 4 * e2 + s2);
    par[a1] = b1;
  }
  // -- provisional ids: corners by class, side interiors per joined pair
  std::vector<int> cid(nside, -1), sfirst(nside, -1);
  std::vector<int> prov(nel * next);
  std::vector<double> gx, gy;
  int ng = 0;
  auto newnode = [&](double px, double py) { gx.push_back(px); gy.push_back(py); return ng++; };
  const std::vector<int>& em = M.ops.emap;
  for (int e = 0; e < nel; ++e)
    for (int s = 0; s < 4; ++s) {
      const int i = 4 * e + s;
      const int root = find_root(par, i);
      if (cid[root] < 0) {
        const int q0 = e * M.npnp + em[s * nm];
        cid[root] = newnode(M.x[q0], M.y[q0]);
      }
      prov[e * next + s * nm] = cid[root];
      const int j = nbr[i];
      if (j >= 0 && j < i) {
        for (int k = 0; k < ni; ++k) prov[e * next + s * nm + 1 + k] = sfirst[j] + (ni - 1 - k);
        sfirst[i] = sfirst[j];
      } else {
        sfirst[i] = ng;
        for (int k = 0; k < ni; ++k) {
          const int q = e * M.npnp + em[s * nm + 1 + k];
          newnode(M.x[q], M.y[q]);
          prov[e * next + s * nm + 1 + k] = sfirst[i] + k;
        }
      }
    }
  // -- essential nodes and the boundary list
  std::vector<char> ess(ng, 0);
  M.belmt.clear(); M.bside.clear(); M.bkind.clear();
  for (int e = 0; e < nel; ++e)
    for (int s = 0; s < 4; ++s) {
      const int k = skind[4 * e + s];
      if (!k) continue;
      if (k != 1 && k != 2) return SEMX_E_ARG;
      M.belmt.push_back(e); M.bside.push_back(s); M.bkind.push_back(k);
      if (k == 1)
        for (int i = 0; i <= nm; ++i) ess[prov[e * next + (s * nm + i) % next]] = 1;
    }
  M.nbound = static_cast<int>(M.belmt.size());
  int nunk = 0;
  for (int v = 0; v < ng; ++v) nunk += !ess[v];
  // -- unknown-node graph (nodes sharing an element)
  std::vector<int> uid(ng, -1), uold;
  for (int v = 0; v < ng; ++v) if (!ess[v]) { uid[v] = (int)uold.size(); uold.push_back(v); }
  Graph G;
  G.n = nunk;
  {
    std::vector<std::vector<int>> nbr(nunk);
    for (int e = 0; e < nel; ++e)
      for (int i = 0; i < next; ++i) {
        const int a = uid[prov[e * next + i]];
        if (a < 0) continue;
        for (int k = 0; k < next; ++k) {
          const int b = uid[prov[e * next + k]];
          if (b >= 0 && b != a) nbr[a].push_back(b);
        }
      }
    G.ptr.assign(nunk + 1, 0);
    for (int a = 0; a < nunk; ++a) {
      std::sort(nbr[a].begin(), nbr[a].end());
      nbr[a].erase(std::unique(nbr[a].begin(), nbr[a].end()), nbr[a].end());
      G.ptr[a + 1] = G.ptr[a] + (int)nbr[a].size();
    }
    G.adj.reserve(G.ptr[nunk]);
    for (int a = 0; a < nunk; ++a) G.adj.insert(G.adj.end(), nbr[a].begin(), nbr[a].end());
  }
  std::vector<int> bmask(nel * next);
  for (int i = 0; i < nel * next; ++i) bmask[i] = ess[prov[i]];
  if (given) {
    const GivenNumbering& gv = *given;
    if (gv.nglobal != ng || gv.nsolve != nunk) return SEMX_E_MESH;
    std::vector<int> to(ng, -1), from(ng, -1);
    for (int i = 0; i < nel * next; ++i) {
      const int g = gv.btog[i], p = prov[i];
      if (g < 0 || g >= ng || gv.bmask[i] != bmask[i]) return SEMX_E_MESH;
      if ((bmask[i] != 0) != (g >= nunk)) return SEMX_E_MESH;
      if ((to[p] >= 0 && to[p] != g) || (from[g] >= 0 && from[g] != p)) return SEMX_E_MESH;
      to[p] = g;
      from[g] = p;
    }
    M.btog.assign(gv.btog, gv.btog + nel * next);
    M.bmask = bmask;
    M.nglobal = ng;
    M.nsolve = nunk;
    M.nband = band_of(M.btog, M.bmask, nel, next);
    M.emask.assign(nel, 0);
    for (int e = 0; e < nel; ++e)
      for (int i = 0; i < next; ++i) if (M.bmask[e * next + i]) M.emask[e] = 1;
    return SEMX_OK;
  }
  if (enum_level >= 0) {
    if (!enumerate_numbering(prov, bmask, ng, nunk, nel, next, enum_level, M.btog))
      return SEMX_E_UNSUPPORTED;
    M.bmask = bmask;
    M.nglobal = ng;
    M.nsolve = nunk;
    M.nband = band_of(M.btog, M.bmask, nel, next);
    M.emask.assign(nel, 0);
    for (int e = 0; e < nel; ++e)
      for (int i = 0; i < next; ++i) if (M.bmask[e * next + i]) M.emask[e] = 1;
    return SEMX_OK;
  }
  // -- candidate orderings; keep the narrowest band (the role of enumerate's
  //    -O2/-O3 root searches, utility/enumerate.cpp:680-809).
  auto apply = [&](const std::vector<int>& perm, std::vector<int>& btog) {
    std::vector<int> newid(ng, -1);
    for (int k = 0; k < nunk; ++k) newid[uold[perm[k]]] = k;
    int ne = 0;
    for (int v = 0; v < ng; ++v) if (ess[v]) newid[v] = nunk + ne++;
    btog.resize(nel * next);
    for (int i = 0; i < nel * next; ++i) btog[i] = newid[prov[i]];
  };
  std::vector<std::vector<int>> cands;
  if (nunk > 0) {
    std::vector<int> natural(nunk);
    std::iota(natural.begin(), natural.end(), 0);
    cands.push_back(natural);
    for (int start : {0, nunk / 3, (2 * nunk) / 3, nunk - 1}) cands.push_back(rcm(G, start));
    // coordinate sorts (row by row), plain and folded (rows taken from both
    // ends alternately, which keeps periodic wrap-around neighbours close)
    for (int axis = 0; axis < 2; ++axis)
      for (int fold = 0; fold < 2; ++fold) {
        const double* c1 = axis ? gx.data() : gy.data();
        const double* c2 = axis ? gy.data() : gx.data();
        double lo = 1e300, hi = -1e300;
        for (int v : uold) { lo = std::min(lo, c1[v]); hi = std::max(hi, c1[v]); }
        const double span = std::max(hi - lo, 1e-300);
        const double eps = 1e-9 * std::max(1.0, span);
        auto key = [&](int v, int& side) {
          const double t = (c1[v] - lo) / span;
          side = (t > 0.5 + 1e-12);
          return fold ? std::min(t, 1.0 - t) * span : c1[v] - lo;
        };
        std::vector<int> p = natural;
        std::stable_sort(p.begin(), p.end(), [&](int a, int b) {
          const int va = uold[a], vb = uold[b];
          int sa, sb;
          const double ka = key(va, sa), kb = key(vb, sb);
          if (std::fabs(ka - kb) > eps) return ka < kb;
          if (fold && sa != sb) return sa < sb;
          return c2[va] < c2[vb];
        });
        cands.push_back(p);
      }
  } else
    cands.push_back({});
  int best = -1, bestband = INT32_MAX;
  std::vector<int> bt;
  for (size_t c = 0; c < cands.size(); ++c) {
    apply(cands[c], bt);
    const int b = band_of(bt, bmask, nel, next);
    if (b < bestband) { bestband = b; best = (int)c; }
  }
  apply(cands[best], M.btog);
  M.bmask = bmask;
  M.nglobal = ng;
  M.nsolve = nunk;
  M.nband = bestband;
  M.emask.assign(nel, 0);
  for (int e = 0; e < nel; ++e)
    for (int i = 0; i < next; ++i) if (M.bmask[e * next + i]) M.emask[e] = 1;
  return SEMX_OK;
}

// ---------------------------------------------------------------------------

// Node coordinates, geometric factors and their deduplicated sets from the
// stored vertices (cylindrical meshes: radius-weighted, src/element.cpp:
// 1281-1436).
static int mesh_geometry(Mesh& M) {
  const int npnp = M.npnp, nel = M.nel;
  const double* vx = M.vx.data();
  const double* vy = M.vy.data();
  M.x.assign((size_t)nel * npnp, 0.0);
  M.y.assign((size_t)nel * npnp, 0.0);
  std::vector<double> g((size_t)nel * G_N * npnp);
  for (int e = 0; e < nel; ++e) {
    element_nodes(M.ops, vx + 4 * e, vy + 4 * e, &M.x[e * npnp], &M.y[e * npnp]);
    // Cartesian factors depend on vertex differences only: translated
    // copies of one element shape give equal factors and share one set.
    if (M.cyl)
      for (int q = 0; q < npnp; ++q)
        if (M.y[e * npnp + q] < 0.0) return SEMX_E_MESH;   // negative radius
    const int st = element_factors(M.ops, vx + 4 * e, vy + 4 * e, &g[(size_t)e * G_N * npnp],
                                   M.cyl ? &M.y[e * npnp] : nullptr);
    if (st) return st;
  }
  // -- deduplicate geometric factor sets (femlib/family.c:171-192 semantics)
  M.eset.assign(nel, -1);
  M.G.clear(); M.setQ3.clear();
  M.nset = 0;
  std::vector<int> reps;
  // Tolerance per factor array: 6e-14 relative to the array's magnitude
  // (semtex compares within 6e-14 absolute, femlib/family.c:171-192, which
  // separates translated copies whose inverse partials are O(1/h)).
  std::vector<double> scale(G_N, 1.0);
  for (int k = 0; k < G_N; ++k)
    for (size_t q = 0; q < (size_t)nel * npnp; ++q) {
      const size_t e = q / npnp, i = q % npnp;
      scale[k] = std::max(scale[k], std::fabs(g[e * G_N * npnp + k * npnp + i]));
    }
  for (int e = 0; e < nel; ++e) {
    const double* ge = &g[(size_t)e * G_N * npnp];
    int found = -1;
    for (int s = 0; s < M.nset && found < 0; ++s) {
      const double* gs = &g[(size_t)reps[s] * G_N * npnp];
      bool same = true;
      for (int q = 0; q < G_N * npnp && same; ++q)
        same = std::fabs(ge[q] - gs[q]) <= EPSDP * scale[q / npnp];
      if (same) found = s;
    }
    if (found < 0) {
      found = M.nset++;
      reps.push_back(e);
      M.G.insert(M.G.end(), ge, ge + G_N * npnp);
      bool q3 = false;
      for (int q = 0; q < npnp; ++q) q3 |= ge[G_Q3 * npnp + q] != 0.0;
      M.setQ3.push_back(q3);
    }
    M.eset[e] = found;
  }
  return SEMX_OK;
}

// Position of node i (0..np-1, in the side's direction) of side s.
static int side_node_index(int np, int s, int i) {
  const int nm = np - 1;
  switch (s) {
    case 0: return i;
    case 1: return i * np + nm;
    case 2: return nm * np + (nm - i);
    default: return (nm - i) * np;
  }
}

// Inverse global mass and side quadrature weights (after the numbering).
static void mesh_weights(Mesh& M) {
  const int np = M.np, npnp = M.npnp, nel = M.nel, next = M.next, nm = np - 1;
  // -- inverse global mass (src/bcmgr.cpp:806-822)
  M.imass.assign(M.nglobal, 0.0);
  for (int e = 0; e < nel; ++e) {
    const double* Q4 = &M.G[(size_t)M.eset[e] * G_N * npnp + G_Q4 * npnp];
    for (int i = 0; i < next; ++i) M.imass[M.btog[e * next + i]] += Q4[M.ops.emap[i]];
  }
  for (double& v : M.imass) v = 1.0 / v;
  // -- side quadrature weights |dx/dxi| w (src/element.cpp:1653-1744)
  M.barea.assign((size_t)M.nbound * np, 0.0);
  for (int b = 0; b < M.nbound; ++b) {
    const int e = M.belmt[b], s = M.bside[b];
    for (int i = 0; i < np; ++i) {
      double tx = 0, ty = 0;
      for (int k = 0; k < np; ++k) {
        int q;
        switch (s) {
          case 0: q = k; break;
          case 1: q = k * np + nm; break;
          case 2: q = nm * np + k; break;
          default: q = k * np; break;
        }
        tx += M.ops.DV[i * np + k] * M.x[e * npnp + q];
        ty += M.ops.DV[i * np + k] * M.y[e * npnp + q];
      }
      M.barea[b * np + i] = std::sqrt(tx * tx + ty * ty) * M.ops.w[i];
      // cylindrical: times the radius (src/edge.cpp:75)
      if (M.cyl) M.barea[b * np + i] *= M.y[e * npnp + side_node_index(np, s, i)];
    }
  }
}

int mesh_set_cylindrical(Mesh& M) {
  if (M.on_device || M.used) return SEMX_E_ARG;
  if (M.cyl) return SEMX_OK;
  M.cyl = true;
  if (int st = mesh_geometry(M)) { M.cyl = false; mesh_geometry(M); return st; }
  mesh_weights(M);
  return SEMX_OK;
}

int mesh_build(Mesh& M, int np, int nel, const double* vx, const double* vy,
               const int* nbr, const int* skind, int svv_mn, double svv_eps,
               double kinvis, const int* btog, const int* bmask, int nglobal, int nsolve,
               int enum_level) {
  if (np < 3 || np > 16 || nel < 1) return SEMX_E_ARG;
  M.np = np; M.nel = nel; M.npnp = np * np;
  M.next = 4 * (np - 1); M.nint = M.npnp - M.next;
  M.nPlane = nel * M.npnp;
  M.planeSize = M.nPlane + (M.nPlane % 2);
  build_ops(M.ops, np, svv_mn, svv_eps, kinvis);
  M.vx.assign(vx, vx + 4 * (size_t)nel);
  M.vy.assign(vy, vy + 4 * (size_t)nel);
  if (int st = mesh_geometry(M)) return st;
  const GivenNumbering given{btog, bmask, nglobal, nsolve};
  int st = number_mesh(M, nbr, skind, (btog && bmask) ? &given : nullptr, enum_level);
  if (st) return st;
  const int next = M.next, nm = np - 1;
  mesh_weights(M);
  // -- essential values: last writer in boundary order (Field::getEssential,
  //    src/field.cpp:994-1019 with Essential::set, condition.cpp:80-101)
  const int nzero = M.nglobal - M.nsolve;
  M.ess_src.assign(nzero, -1);
  for (int b = 0; b < M.nbound; ++b) {
    if (M.bkind[b] != 1) continue;
    const int* bmap = &M.btog[M.belmt[b] * next];
    const int s = M.bside[b];
    for (int i = 0; i <= nm; ++i) {
      const int gid = bmap[(s * nm + i) % next];
      M.ess_src[gid - M.nsolve] = b * np + i;
    }
  }
  for (int v : M.ess_src) if (v < 0) return SEMX_E_MESH;
  // -- assembly lists: global node -> element-boundary slots (owner gather,
  //    replaces the scatter-sum of Veclib::scatr_sum)
  M.asm_ptr.assign(M.nglobal + 1, 0);
  for (int i = 0; i < nel * next; ++i) M.asm_ptr[M.btog[i] + 1]++;
  for (int gI = 0; gI < M.nglobal; ++gI) M.asm_ptr[gI + 1] += M.asm_ptr[gI];
  M.asm_idx.assign(nel * next, 0);
  {
    std::vector<int> fill(M.asm_ptr.begin(), M.asm_ptr.end() - 1);
    for (int i = 0; i < nel * next; ++i) M.asm_idx[fill[M.btog[i]]++] = i;
  }
  // -- natural boundaries per element
  M.nat_ptr.assign(nel + 1, 0);
  for (int b = 0; b < M.nbound; ++b) if (M.bkind[b] == 2) M.nat_ptr[M.belmt[b] + 1]++;
  for (int e = 0; e < nel; ++e) M.nat_ptr[e + 1] += M.nat_ptr[e];
  M.nat_b.assign(M.nat_ptr[nel], 0);
  {
    std::vector<int> fill(M.nat_ptr.begin(), M.nat_ptr.end() - 1);
    for (int b = 0; b < M.nbound; ++b)
      if (M.bkind[b] == 2) M.nat_b[fill[M.belmt[b]]++] = b;
  }
  return SEMX_OK;
}

// ---------------------------------------------------------------------------
// Elemental Helmholtz operator and static condensation.
// ---------------------------------------------------------------------------

// y = (K + hcon M) x on one element by sum factorisation, the operator of
// src/element.cpp:1564-1650; dv/dt the (SVV) derivative and transpose.
// q4: the mass factors (null: the set's Q4; cylindrical meshes pass the
// mode's radius-weighted factors with hcon = 1).
static void elmt_apply(int np, const double* g, const double* dv, const double* dt,
                       double hcon, const double* x, double* y, const double* q4 = nullptr) {
  const int npnp = np * np;
  std::vector<double> R(npnp), S(npnp);
  const double *Q1 = g + G_Q1 * npnp, *Q2 = g + G_Q2 * npnp, *Q3 = g + G_Q3 * npnp,
               *Q4 = q4 ? q4 : g + G_Q4 * npnp;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      double r = 0, s = 0;
      for (int q = 0; q < np; ++q) {
        r += x[i * np + q] * dt[q * np + j];
        s += dv[i * np + q] * x[q * np + j];
      }
      const int k = i * np + j;
      R[k] = Q1[k] * r + Q3[k] * s;
      S[k] = Q2[k] * s + Q3[k] * r;
    }
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      double t = Q4[i * np + j] * x[i * np + j] * hcon;
      for (int q = 0; q < np; ++q) t += dt[i * np + q] * S[q * np + j] + R[i * np + q] * dv[q * np + j];
      y[i * np + j] = t;
    }
}

// Cholesky of a small dense SPD matrix in place (lower), returns 0 on success.
static int dense_chol(int n, double* A) {
  for (int j = 0; j < n; ++j) {
    double d = A[j * n + j];
    for (int k = 0; k < j; ++k) d -= A[j * n + k] * A[j * n + k];
    if (!(d > 0.0)) return 1;
    d = std::sqrt(d);
    A[j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double s = A[i * n + j];
      for (int k = 0; k < j; ++k) s -= A[i * n + k] * A[j * n + k];
      A[i * n + j] = s / d;
    }
  }
  return 0;
}

// Inverse of an SPD matrix from its Cholesky factor (full symmetric result).
static int spd_inverse(int n, double* A) {
  if (dense_chol(n, A)) return 1;
  std::vector<double> Linv(n * n, 0.0);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) s -= A[i * n + k] * Linv[k * n + c];
      Linv[i * n + c] = (i < c) ? 0.0 : s / A[i * n + i];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = std::max(i, j); k < n; ++k) s += Linv[k * n + i] * Linv[k * n + j];
      A[i * n + j] = s;
    }
  return 0;
}

// Elemental matrix in emap order, condensed: hbb <- hbb - hbi hii^-1 hib,
// hbi <- hbi hii^-1, hii <- hii^-1 (src/element.cpp:131-229).
static int condense(const Mesh& M, const double* g, double lambda2, double betak2,
                    double* hbb, double* hbi, double* hii, const double* q4 = nullptr) {
  const int np = M.np, npnp = M.npnp, next = M.next, nint = M.nint;
  const double* dv = (lambda2 > EPSDP) ? M.ops.SDV.data() : M.ops.DV.data();
  const double* dt = (lambda2 > EPSDP) ? M.ops.SDT.data() : M.ops.DT.data();
  std::vector<double> H((size_t)npnp * npnp), e(npnp), col(npnp);
  for (int c = 0; c < npnp; ++c) {         // column c (emap order) = H e_c
    std::fill(e.begin(), e.end(), 0.0);
    e[M.ops.emap[c]] = 1.0;
    elmt_apply(np, g, dv, dt, q4 ? 1.0 : lambda2 + betak2, e.data(), col.data(), q4);
    for (int r = 0; r < npnp; ++r) H[(size_t)r * npnp + c] = col[M.ops.emap[r]];
  }
  // symmetrise (the operator is symmetric; removes roundoff asymmetry)
  for (int r = 0; r < npnp; ++r)
    for (int c = r + 1; c < npnp; ++c) {
      const double a = 0.5 * (H[(size_t)r * npnp + c] + H[(size_t)c * npnp + r]);
      H[(size_t)r * npnp + c] = H[(size_t)c * npnp + r] = a;
    }
  for (int r = 0; r < next; ++r)
    for (int c = 0; c < next; ++c) hbb[r * next + c] = H[(size_t)r * npnp + c];
  if (nint == 0) return SEMX_OK;
  for (int r = 0; r < nint; ++r)
    for (int c = 0; c < nint; ++c) hii[r * nint + c] = H[(size_t)(next + r) * npnp + next + c];
  if (spd_inverse(nint, hii)) return SEMX_E_NOT_SPD;
  std::vector<double> hb(next * nint);
  for (int r = 0; r < next; ++r)
    for (int c = 0; c < nint; ++c) hb[r * nint + c] = H[(size_t)r * npnp + next + c];
  for (int r = 0; r < next; ++r)
    for (int c = 0; c < nint; ++c) {
      double s = 0;
      for (int k = 0; k < nint; ++k) s += hb[r * nint + k] * hii[k * nint + c];
      hbi[r * nint + c] = s;
    }
  for (int r = 0; r < next; ++r)
    for (int c = 0; c < next; ++c) {
      double s = 0;
      for (int k = 0; k < nint; ++k) s += hbi[r * nint + k] * hb[c * nint + k];
      hbb[r * next + c] -= s;
    }
  return SEMX_OK;
}

// Inverse of the transposed diagonal block of the band factor at rows
// [j0, j0+nb): L = (U_JJ^T)^-1 (lower triangular, identity beyond n) and
// LT = L^T.
static void block_inverse(const double* A, int n, int kd, int ldab, int nb, int j0,
                          double* L, double* LT) {
  std::vector<double> T((size_t)nb * nb, 0.0);  // T = U_JJ^T (lower), identity pad
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b <= a; ++b) {
      const int r = j0 + b, c = j0 + a;
      if (c < n) T[a * nb + b] = (c - r <= kd) ? A[(size_t)r * ldab + (c - r)] : 0.0;
      else T[a * nb + b] = (a == b) ? 1.0 : 0.0;
    }
  for (int col = 0; col < nb; ++col)
    for (int a = 0; a < nb; ++a) {
      double s = (a == col) ? 1.0 : 0.0;
      for (int b = col; b < a; ++b) s -= T[a * nb + b] * L[b * nb + col];
      L[a * nb + col] = (a < col) ? 0.0 : s / T[a * nb + a];
    }
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b) LT[b * nb + a] = L[a * nb + b];
}

// Banded Cholesky on row band storage A[i*ldab + d] = A(i, i+d), d = 0..kd.
static int band_chol_rows(int n, int kd, int ldab, double* A) {
  for (int j = 0; j < n; ++j) {
    double* rj = A + (size_t)j * ldab;
    if (!(rj[0] > 0.0)) return SEMX_E_NOT_SPD;
    const double ujj = std::sqrt(rj[0]);
    rj[0] = ujj;
    const int km = std::min(kd, n - 1 - j);
    const double inv = 1.0 / ujj;
    for (int d = 1; d <= km; ++d) rj[d] *= inv;
    for (int i = 1; i <= km; ++i) {
      const double u = rj[i];
      if (u == 0.0) continue;
      double* ri = A + (size_t)(j + i) * ldab;
      const int len = km - i;
      for (int l = 0; l <= len; ++l) ri[l] -= u * rj[i + l];
    }
  }
  return SEMX_OK;
}

// Spikes of the partitioned band sweep for mode m from its factor
// (row band storage A[j*ldab + d] = U(j, j+d)).  With L = U^T, partition p
// holding rows [s, e):
//   forward  V_p = L_p^-1 [B_p; 0], B_p(r, c) = L(s+r, s-kd+c), r < kd:
//            the solution of L y = b is y_p = L_p^-1 b_p - V_p y[s-kd, s);
//   backward W_p = U_p^-1 [0; C_p], C_p(r, c) = U(s+r, e+c):
//            the solution of U x = y is x_p = U_p^-1 y_p - W_p x[e, e+kd).
// Each by substitution over the partition's rows with all kd columns at
// once; stored column-major, each column n_p rows padded to an even count.
// (cloned for AVX2+FMA hosts: the kd-wide row updates are most of the
// setup time of a partitioned system)
__attribute__((target_clones("arch=haswell", "default")))
static void band_spikes(Helm& H, int m, const double* A) {
  const int P = H.npart, kd = H.kd, ldab = H.ldab;
  const int* pr = &H.prow[(size_t)m * (P + 1)];
  double* out = &H.spk[H.spk_off[m]];
  std::vector<double> T;
  auto store = [&](int np_) {   // columns of an even number of rows (16-byte aligned)
    const int ld = (np_ + 1) & ~1;
    for (int r = 0; r < np_; ++r)
      for (int c = 0; c < kd; ++c) out[(size_t)c * ld + r] = T[(size_t)r * kd + c];
    out += (size_t)ld * kd;
  };
  for (int p = 1; p < P; ++p) {
    const int s = pr[p], np_ = pr[p + 1] - s;
    T.assign((size_t)np_ * kd, 0.0);
    for (int r = 0; r < np_; ++r) {
      const int g = s + r;
      double* row = &T[(size_t)r * kd];
      for (int c = r; c < kd; ++c) {
        const int h = s - kd + c;
        row[c] = A[(size_t)h * ldab + (g - h)];
      }
      for (int k = std::max(0, r - kd); k < r; ++k) {
        const double l = A[(size_t)(s + k) * ldab + (r - k)];
        if (l == 0.0) continue;
        const double* vk = &T[(size_t)k * kd];
        for (int c = 0; c < kd; ++c) row[c] -= l * vk[c];
      }
      const double d = A[(size_t)g * ldab];
      for (int c = 0; c < kd; ++c) row[c] /= d;
    }
    store(np_);
  }
  for (int p = 0; p + 1 < P; ++p) {
    const int s = pr[p], e = pr[p + 1], np_ = e - s;
    T.assign((size_t)np_ * kd, 0.0);
    for (int r = np_ - 1; r >= 0; --r) {
      const int g = s + r;
      double* row = &T[(size_t)r * kd];
      for (int c = 0; c < kd && e + c - g <= kd; ++c) row[c] = A[(size_t)g * ldab + (e + c - g)];
      for (int k = r + 1; k <= std::min(np_ - 1, r + kd); ++k) {
        const double u = A[(size_t)g * ldab + (k - r)];
        if (u == 0.0) continue;
        const double* wk = &T[(size_t)k * kd];
        for (int c = 0; c < kd; ++c) row[c] -= u * wk[c];
      }
      const double d = A[(size_t)g * ldab];
      for (int c = 0; c < kd; ++c) row[c] /= d;
    }
    store(np_);
  }
}

int helm_build(Helm& H) {
  Mesh& M = *H.M;
  M.used = true;
  const int next = M.next, nint = M.nint, nset = M.nset, npnp = M.npnp;
  const int nmodes = H.nmodes;
  H.npts = M.nglobal + M.nel * nint;
  H.kd = M.nband - 1;
  H.ldab = M.nband;
  H.modes.assign(nmodes, ModeSys());
  H.hbb.assign((size_t)nmodes * nset * next * next, 0.0);
  H.hbi.assign((size_t)nmodes * nset * next * std::max(nint, 1), 0.0);
  H.hii.assign((size_t)nmodes * nset * std::max(nint * nint, 1), 0.0);
  // Band-solve blocking: NB columns per step, the off-diagonal entries of a
  // block stored as a zero-padded [NB][kdp] panel, kdp = kd rounded to 64.
  {
    const int kd = std::max(H.kd, 1);
    // thread r owns update row r; bands wider than 512 use narrower blocks
    // and several rows per thread
    // cooperative band solve (several workgroups per mode): by default for
    // bands wider than 1008 (narrower ones stream extended panels faster
    // with one workgroup per mode); SEMX_BAND_COOP=0 / 1 forces it off / on
    const char* env = getenv("SEMX_BAND_COOP");
    H.coop_mode = env ? atoi(env) != 0 : kd > 1008;
    if (kd <= 512) { H.nb = 32; H.npass = 1; H.bt = 64; while (H.bt < kd) H.bt *= 2; }
    else if (kd <= 1024) { H.nb = 16; H.npass = 2; H.bt = 512; }
    else if (kd <= 2048) { H.nb = 8; H.npass = 4; H.bt = 512; }
    else { H.nb = 8; H.npass = 0; H.bt = 0; }   // no direct solve for this band
    // bands up to 1008 wide solve with the inverses folded into the panels
    // (32-column blocks up to 512, 16-column blocks beyond: one thread per
    // panel row, at most 1024)
    // SEMX_BAND_XCOOP=1: the cooperative sweep over extended panels
    const char* envx = getenv("SEMX_BAND_XCOOP");
    H.xcoop = envx && atoi(envx) != 0 && kd <= 1008;
    if (H.xcoop) H.coop_mode = true;
    H.band_ext = kd <= 1008 && (!H.coop_mode || H.xcoop);
    H.ldx = (H.nb + kd + 7) / 8 * 8;
    H.kdp = (kd + 63) / 64 * 64;
  }
  size_t offH = 0, offL = 0;
  const int nb = H.nb;
  // SVV weights of the Fourier modes (SVV::coeffs_z, src/svv.cpp:82-123:
  // N = nMode - 1, nMode the field's Fourier mode count, src/geometry.h:42)
  std::vector<double> Sz(nmodes, 1.0);
  {
    const int N = H.svv_nmode - 1, mz = H.svv_mz;
    if (mz >= 0 && mz < N && H.svv_epsz > EPSDP)
      for (int m = 0; m < nmodes; ++m) {
        const int i = H.base + m;
        const double q = (i > mz) ? std::exp(-std::pow((N - i) / (1.0 * mz - i), 2)) : 0.0;
        Sz[m] = 1.0 + H.svv_epsz / H.svv_nu * q;
      }
  }
  // Field::modeConstant (src/field.cpp:1121-1146)
  auto mode_constant = [&](int mode) {
    if (!M.cyl || (H.field != 'v' && H.field != 'w')) return H.beta * mode;
    if (mode == 0) return 1.0;
    return H.field == 'v' ? H.beta * mode + 1.0 : H.beta * mode - 1.0;
  };
  for (int m = 0; m < nmodes; ++m) {
    ModeSys& S = H.modes[m];
    const double bk = mode_constant(H.base + m);
    S.betak2 = bk * bk;
    if (H.lambda2 > EPSDP) S.betak2 *= Sz[m];   // src/matrix.cpp:82-83
    S.kind = (H.method == SEMX_MIXED) ? ((H.base + m == 0) ? SEMX_DIRECT : SEMX_JACPCG)
                                      : H.method;
    // Singular: no essential BCs and lambda2 + betak2 ~ 0 (matrix.cpp:167-169).
    S.singular = (H.lambda2 + S.betak2) < EPSSP && M.nglobal == M.nsolve;
    S.n = S.singular ? M.nsolve - 1 : M.nsolve;
    if (S.kind == SEMX_DIRECT) {
      if (H.npass == 0) return SEMX_E_UNSUPPORTED;
      S.nblk = (S.n + nb - 1) / nb;
      S.off_H = offH;
      S.off_L = offL;
      offH += (size_t)S.nblk * nb * (H.band_ext ? H.ldx : H.kdp);
      if (!H.band_ext) offL += (size_t)S.nblk * nb * nb;
    }
  }
  H.band_doubles = offH;
  H.inv_doubles = offL;
  // Partitioned sweep: enough partitions that one workgroup per direct
  // mode, plane and partition fills the chip's 256 CUs (at most 8), each
  // partition at least kd + one block of rows (its coupling reaches only
  // the neighbouring partition).  SEMX_BAND_PARTS sets the count (1: off).
  H.npart = 1;
  H.prow.clear(); H.spk_off.clear(); H.spk.clear();
  if (H.band_ext && nb == 32 && H.kd >= 1) {
    int ndirect = 0, nmin = INT32_MAX;
    for (int m = 0; m < nmodes; ++m)
      if (H.modes[m].kind == SEMX_DIRECT) { ++ndirect; nmin = std::min(nmin, H.modes[m].n); }
    int P = ndirect ? std::max(1, std::min(8, 256 / (2 * ndirect))) : 1;
    if (const char* ep = getenv("SEMX_BAND_PARTS")) P = std::max(1, std::min(8, atoi(ep)));
    if (ndirect) P = std::min(P, nmin / (H.kd + nb));
    if (P >= 2) {
      H.npart = P;
      H.prow.assign((size_t)nmodes * (P + 1), 0);
      H.spk_off.assign(nmodes, 0);
      size_t off = 0;
      bool fits = true;
      for (int m = 0; m < nmodes; ++m) {
        const ModeSys& S = H.modes[m];
        if (S.kind != SEMX_DIRECT) continue;
        int* pr = &H.prow[(size_t)m * (P + 1)];
        for (int q = 0; q < P; ++q) pr[q] = (int)((long long)q * S.nblk / P) * nb;
        pr[P] = S.n;
        for (int q = 0; q < P; ++q) fits &= pr[q + 1] - pr[q] >= H.kd + 1;
        H.spk_off[m] = (long long)off;
        auto ldq = [&](int q) { return (size_t)((pr[q + 1] - pr[q] + 1) & ~1); };
        for (int q = 1; q < P; ++q) off += ldq(q) * H.kd;
        for (int q = 0; q + 1 < P; ++q) off += ldq(q) * H.kd;
      }
      if (!fits) {
        H.npart = 1;
        H.prow.clear(); H.spk_off.clear();
      } else {
        H.spk.assign(off, 0.0);
      }
    }
  }
  if (H.band_ext) {
    H.xblk = 0;
    for (int m = 0; m < nmodes; ++m)
      if (H.modes[m].kind == SEMX_DIRECT) H.xblk = std::max(H.xblk, H.modes[m].nblk);
    H.xrows.assign((size_t)nmodes * std::max(H.xblk, 1) * 2, nb);
  }
  H.Pf.assign(offH, 0.0);
  H.Pb.assign(offH, 0.0);
  H.Linv.assign(offL, 0.0);
  H.LinvT.assign(offL, 0.0);
  H.pc.assign((size_t)nmodes * H.npts, 0.0);

  // cylindrical meshes: the mass term of each mode, Q4 y (lambda2 + betak2 /
  // y^2) per node, zero on the axis (Element::HelmholtzKern and
  // HelmholtzDiag, src/element.cpp:1512-1599: hCon = betak2/r^2 + lambda2
  // where r^2 > EPSDP, else 0)
  H.mq.clear();
  if (M.cyl) {
    H.mq.assign((size_t)nmodes * nset * npnp, 0.0);
    for (int m = 0; m < nmodes; ++m)
      for (int s = 0; s < nset; ++s) {
        const double* g = &M.G[(size_t)s * G_N * npnp];
        double* q = &H.mq[((size_t)m * nset + s) * npnp];
        for (int i = 0; i < npnp; ++i) {
          const double y = g[G_Y * npnp + i], r2 = y * y;
          const double hc = (r2 > EPSDP) ? H.modes[m].betak2 / r2 + H.lambda2 : 0.0;
          q[i] = g[G_Q4 * npnp + i] * y * hc;
        }
      }
  }
  auto mass_of = [&](int m, int s) -> const double* {
    return M.cyl ? &H.mq[((size_t)m * nset + s) * npnp] : nullptr;
  };
  std::vector<int> status(nmodes, SEMX_OK);
  auto work = [&](int m) {
    ModeSys& S = H.modes[m];
    // elemental matrices per geometric set
    for (int s = 0; s < nset; ++s) {
      const double* g = &M.G[(size_t)s * G_N * npnp];
      const size_t ms = (size_t)m * nset + s;
      const int st = condense(M, g, H.lambda2, S.betak2, &H.hbb[ms * next * next],
                              &H.hbi[ms * next * std::max(nint, 1)],
                              &H.hii[ms * std::max(nint * nint, 1)], mass_of(m, s));
      if (st) { status[m] = st; return; }
    }
    if (S.kind == SEMX_DIRECT) {
      const int n = S.n, kd = H.kd, ldab = H.ldab, kdp = H.kdp;
      std::vector<double> band((size_t)std::max(n, 1) * ldab, 0.0);
      double* A = band.data();
      // assemble condensed boundary matrices into row band storage
      // A[r*ldab + (c-r)] = A(r, c), c >= r (src/matrix.cpp:236-241)
      for (int e = 0; e < M.nel; ++e) {
        const double* hbb = &H.hbb[((size_t)m * nset + M.eset[e]) * next * next];
        const int* bm = &M.btog[e * next];
        for (int i = 0; i < next; ++i) {
          const int r = bm[i];
          if (r >= n) continue;
          for (int k = 0; k < next; ++k) {
            const int c = bm[k];
            if (c >= n || c < r) continue;
            A[(size_t)r * ldab + (c - r)] += hbb[i * next + k];
          }
        }
      }
      const int st = band_chol_rows(n, kd, ldab, A);
      if (st) { status[m] = st; return; }
      // forward panels: Pf[J][jj][r] = U(j0+jj, j0+nb+r); backward panels:
      // Pb[J][jj][r] = U(j0-1-r, j0+jj); zero outside the band.  With
      // extended panels they are built per block in scratch and folded below.
      const bool ext = H.band_ext;
      const int ldx = H.ldx;
      std::vector<double> pf_s, pb_s, L_s, LT_s;
      if (ext) {
        pf_s.assign((size_t)nb * kdp, 0.0);
        pb_s.assign((size_t)nb * kdp, 0.0);
        L_s.assign((size_t)nb * nb, 0.0);
        LT_s.assign((size_t)nb * nb, 0.0);
      }
      for (int J = 0; J < S.nblk; ++J) {
        const int j0 = J * nb;
        double* pf = ext ? pf_s.data() : &H.Pf[S.off_H + (size_t)J * nb * kdp];
        double* pb = ext ? pb_s.data() : &H.Pb[S.off_H + (size_t)J * nb * kdp];
        if (ext) {
          std::fill(pf_s.begin(), pf_s.end(), 0.0);
          std::fill(pb_s.begin(), pb_s.end(), 0.0);
        }
        for (int jj = 0; jj < nb; ++jj) {
          const int j = j0 + jj;
          if (j >= n) continue;
          for (int r = 0; r < kdp; ++r) {
            const int df = nb + r - jj, i_f = j0 + nb + r;
            if (df <= kd && i_f < n) pf[(size_t)jj * kdp + r] = A[(size_t)j * ldab + df];
            const int db = jj + 1 + r, i_b = j0 - 1 - r;
            if (db <= kd && i_b >= 0) pb[(size_t)jj * kdp + r] = A[(size_t)i_b * ldab + db];
          }
        }
        if (!ext) continue;
        // extended panels of block J, column kk of the window:
        //   Xf[kk][r]      = Linv[r][kk]                      r < nb
        //   Xf[kk][nb+r]   = sum_jj Pf[jj][r] Linv[jj][kk]    r < kd
        // and the same with Pb and LinvT for the backward sweep
        block_inverse(A, n, kd, ldab, nb, j0, L_s.data(), LT_s.data());
        // stored as column pairs: entry (kk, r) at ((kk/2)*ldx + r)*2 + kk%2,
        // so one 16-byte load gives a row its two values of a pair
        double* xf = &H.Pf[S.off_H + (size_t)J * nb * ldx];
        double* xb = &H.Pb[S.off_H + (size_t)J * nb * ldx];
        auto at = [&](int kk, int r) { return ((size_t)(kk >> 1) * ldx + r) * 2 + (kk & 1); };
        for (int kk = 0; kk < nb; ++kk) {
          for (int r = 0; r < nb; ++r) {
            xf[at(kk, r)] = L_s[r * nb + kk];
            xb[at(kk, r)] = LT_s[r * nb + kk];
          }
          for (int r = 0; r < kd; ++r) {
            long double sf = 0.0L, sb = 0.0L;
            for (int jj = 0; jj < nb; ++jj) {
              sf += (long double)pf[(size_t)jj * kdp + r] * L_s[jj * nb + kk];
              sb += (long double)pb[(size_t)jj * kdp + r] * LT_s[jj * nb + kk];
            }
            xf[at(kk, nb + r)] = (double)sf;
            xb[at(kk, nb + r)] = (double)sb;
          }
        }
        // live rows: beyond the block's envelope the folded rows are exact
        // zeros (the factor keeps the matrix's envelope)
        int hf = 0, hb = 0;
        for (int r = 0; r < kd; ++r)
          for (int kk = 0; kk < nb; ++kk) {
            if (xf[at(kk, nb + r)] != 0.0) hf = r + 1;
            if (xb[at(kk, nb + r)] != 0.0) hb = r + 1;
          }
        H.xrows[((size_t)m * H.xblk + J) * 2] = nb + hf;
        H.xrows[((size_t)m * H.xblk + J) * 2 + 1] = nb + hb;
      }
      if (H.npart > 1) band_spikes(H, m, A);
      // diagonal-block inverses: Linv_J = (U_JJ^T)^-1, lower triangular
      for (int J = 0; !ext && J < S.nblk; ++J) {
        block_inverse(A, n, kd, ldab, nb, J * nb, &H.Linv[S.off_L + (size_t)J * nb * nb],
                      &H.LinvT[S.off_L + (size_t)J * nb * nb]);
      }
    } else {
      // Jacobi preconditioner: inverse assembled diagonal (matrix.cpp:276-317)
      double* PC = &H.pc[(size_t)m * H.npts];
      // element diagonal (emap order) per geometric set: diagonal entries of
      // the operator applied to unit vectors (src/element.cpp:1512-1561)
      const double* dv = (H.lambda2 > EPSDP) ? M.ops.SDV.data() : M.ops.DV.data();
      const double* dt = (H.lambda2 > EPSDP) ? M.ops.SDT.data() : M.ops.DT.data();
      std::vector<double> dg((size_t)nset * npnp), ev(npnp), col(npnp);
      for (int s = 0; s < nset; ++s) {
        const double* g = &M.G[(size_t)s * G_N * npnp];
        for (int c = 0; c < npnp; ++c) {
          std::fill(ev.begin(), ev.end(), 0.0);
          ev[M.ops.emap[c]] = 1.0;
          elmt_apply(M.np, g, dv, dt, M.cyl ? 1.0 : H.lambda2 + S.betak2, ev.data(),
                     col.data(), mass_of(m, s));
          dg[(size_t)s * npnp + c] = col[M.ops.emap[c]];
        }
      }
      for (int e = 0; e < M.nel; ++e) {
        const double* d = &dg[(size_t)M.eset[e] * npnp];
        for (int c = 0; c < next; ++c) PC[M.btog[e * next + c]] += d[c];
        for (int c = next; c < npnp; ++c) PC[M.nglobal + e * nint + (c - next)] = d[c];
      }
      for (int i = 0; i < H.npts; ++i) PC[i] = 1.0 / PC[i];
    }
  };
  const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; ++t)
    pool.emplace_back([&, t]() { for (int m = t; m < nmodes; m += nth) work(m); });
  for (auto& th : pool) th.join();
  for (int st : status) if (st) return st;
  // plane list: every plane except the Nyquist plane 1, which only the
  // planes that start at mode 0 hold (ROOTONLY skip, field.cpp:566); with
  // base_mode > 0 plane 1 is Im of mode base_mode and is solved
  H.plane_of.clear();
  for (int k = 0; k < H.nz; ++k)
    if (!(H.nz > 1 && k == 1 && H.base == 0)) H.plane_of.push_back(k);
  H.nplanes = (int)H.plane_of.size();
  H.solstride = M.nglobal + (M.nglobal % 2);
  // band-solve work list: one workgroup per direct mode, solving every
  // plane of that mode (Re and Im share the factor)
  H.mode_plane.assign(2 * nmodes, -1);
  for (int p = 0; p < H.nplanes; ++p) {
    const int k = H.plane_of[p];
    H.mode_plane[2 * (k >> 1) + (k & 1)] = p;
  }
  H.dmode.clear();
  for (int m = 0; m < nmodes; ++m)
    if (H.modes[m].kind == SEMX_DIRECT) H.dmode.push_back(m);
  // PCG work list: one workgroup per plane of a JACPCG mode
  H.pcg_plane.clear();
  for (int p = 0; p < H.nplanes; ++p)
    if (H.modes[H.plane_of[p] >> 1].kind == SEMX_JACPCG) H.pcg_plane.push_back(p);
  H.workstride = 5 * (size_t)H.npts + (size_t)M.nel * next;
  H.workstride += H.workstride & 1;
  return SEMX_OK;
}

}  // namespace semx
