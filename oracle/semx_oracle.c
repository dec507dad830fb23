/*
 * semx_oracle.c -- CPU FP64 restatement of semtex's Helmholtz / DNS hot
 * path.  TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); the
 * product never links it.  See semx_oracle.h for conventions.
 *
 * Each routine cites the reference file:line whose algorithm it restates.
 * Where semtex calls third-party BLAS/LAPACK (dpbtrf, dpbtrs, dgetrf,
 * dgetri, dgemv) the oracle carries its own textbook implementation of the
 * same factorisation, so results agree to roundoff, not bitwise.
 */
#include "semx_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

#define EPSDP 6.0e-14            /* veclib/cveclib.h:58 */
#define EPSSP 6.0e-7             /* veclib/cveclib.h:54 */

static void* xcalloc (size_t n, size_t s)
{
  void* p = calloc (n ? n : 1, s);
  if (!p) { fprintf (stderr, "semx_oracle: out of memory\n"); abort (); }
  return p;
}

/* ======================================================================== *
 * Quadrature and 1D operators.
 * ======================================================================== */

/* Legendre polynomial L_n and derivative at x by the three-term recurrence. */
static void legendre (int n, long double x, long double* p, long double* dp)
{
  long double p0 = 1.0L, p1 = x, d0 = 0.0L, d1 = 1.0L, p2, d2;
  int k;
  if (n == 0) { *p = 1.0L; *dp = 0.0L; return; }
  for (k = 2; k <= n; k++) {
    p2 = ((2*k - 1) * x * p1 - (k - 1) * p0) / k;
    d2 = d0 + (2*k - 1) * p1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  *p = p1; *dp = d1;
}

void orc_zwglj (double* z, double* w, int np)
/* Gauss-Lobatto-Legendre points and weights (alpha = beta = 0), as
 * femlib/polylib.c:312-340 (zwglj): z_0 = -1, z_N = 1, interior points the
 * zeros of L'_N; w_j = 2 / (N (N+1) L_N(z_j)^2).  Interior zeros by Newton
 * iteration in extended precision, made exactly antisymmetric.            */
{
  const int N = np - 1;
  int j, it;
  if (np == 1) { z[0] = 0.0; w[0] = 2.0; return; }
  z[0] = -1.0; z[N] = 1.0;
  for (j = 1; j <= N / 2; j++) {
    long double x = -cosl (3.14159265358979323846264338327950288L * j / N);
    for (it = 0; it < 100; it++) {
      long double p, dp, d2, dx;
      legendre (N, x, &p, &dp);
      d2 = (2.0L * x * dp - N * (N + 1.0L) * p) / (1.0L - x * x);
      dx = dp / d2;
      x -= dx;
      if (fabsl (dx) < 1e-19L) break;
    }
    z[j]     = (double) x;
    z[N - j] = (double) -x;
  }
  if (N % 2 == 0) z[N / 2] = 0.0;
  for (j = 0; j <= N; j++) {
    long double p, dp;
    legendre (N, z[j], &p, &dp);
    w[j] = (double) (2.0L / (N * (N + 1.0L) * p * p));
  }
}

void orc_dglj (double* D, double* DT, const double* z, int np)
/* Derivative matrix of the Lagrange interpolants through the GLL points,
 * femlib/polylib.c:513-556 (Dglj): D[i][j] = h'_j(z_i),
 * D_ij = L_N(z_i) / (L_N(z_j) (z_i - z_j)), D_00 = -N(N+1)/4,
 * D_NN = N(N+1)/4, other diagonal entries zero.  DT is the transpose.    */
{
  const int N = np - 1;
  int i, j;
  long double* L = (long double*) xcalloc (np, sizeof (long double));
  for (i = 0; i < np; i++) { long double dp; legendre (N, z[i], &L[i], &dp); }
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) {
      double d;
      if (i != j)            d = (double) (L[i] / (L[j] * ((long double) z[i] - z[j])));
      else if (i == 0)       d = -0.25 * N * (N + 1);
      else if (i == N)       d =  0.25 * N * (N + 1);
      else                   d = 0.0;
      D [i*np + j] = d;
      DT[j*np + i] = d;
    }
  free (L);
}

void orc_legtran (int np, double* FW, double* BW)
/* GLL Legendre transform matrices, femlib/operators.c:391-480 (dglldpt)
 * with the table of femlib/operators.c:317-374 (dglldpc):
 *   FW[i][j] = C_i w_j L_i(z_j),  BW[i][j] = L_j(z_i),
 *   C_i = (2i+1)/2 for i < N, N/2 for i = N.                             */
{
  const int N = np - 1;
  int i, j;
  double* z = (double*) xcalloc (np, sizeof (double));
  double* w = (double*) xcalloc (np, sizeof (double));
  double* tab = (double*) xcalloc (np * np, sizeof (double));
  orc_zwglj (z, w, np);
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) {
      long double p, dp;
      legendre (i, z[j], &p, &dp);
      tab[i*np + j] = (double) p;
    }
  for (i = 0; i < np; i++) {
    const double ci = (i < N) ? 0.5 * (i + i + 1) : 0.5 * N;
    for (j = 0; j < np; j++) {
      FW[i*np + j] = ci * w[j] * tab[i*np + j];
      BW[i*np + j] = tab[j*np + i];
    }
  }
  free (z); free (w); free (tab);
}

void orc_svv_operators (int np, int svv_mn, double svv_eps, double kinvis,
                        double* SDV, double* SDT)
/* SVV-stabilised derivative operator, src/svv.cpp:126-185:
 *   SDV = BW . diag(sqrt(S)) . FW . DV,   S from src/svv.cpp:41-79 (equal to
 *   1 when SVV_MN < 0, the default femlib/defaults.h:53-54).              */
{
  const int N = np - 1;
  int i, j, k;
  double *z = xcalloc (np, sizeof (double)), *w = xcalloc (np, sizeof (double));
  double *DV = xcalloc (np*np, sizeof (double)), *DT = xcalloc (np*np, sizeof (double));
  double *FW = xcalloc (np*np, sizeof (double)), *BW = xcalloc (np*np, sizeof (double));
  double *S = xcalloc (np, sizeof (double)), *dv = xcalloc (np*np, sizeof (double));
  double *dt = xcalloc (np*np, sizeof (double));

  orc_zwglj (z, w, np);
  orc_dglj (DV, DT, z, np);
  orc_legtran (np, FW, BW);

  if (svv_mn >= 0 && svv_mn < N && svv_eps > EPSDP) {
    for (i = svv_mn + 1; i < np; i++) {
      const double q = (N - i) / (1.0 * svv_mn - i);
      S[i] = exp (-q * q);
    }
    for (i = 0; i < np; i++) S[i] = 1.0 + svv_eps / kinvis * S[i];
  } else
    for (i = 0; i < np; i++) S[i] = 1.0;

  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) dv[i*np + j] = sqrt (S[i]) * FW[i*np + j];
  for (i = 0; i < np; i++)          /* dt = BW . dv */
    for (j = 0; j < np; j++) {
      double s = 0.0;
      for (k = 0; k < np; k++) s += BW[i*np + k] * dv[k*np + j];
      dt[i*np + j] = s;
    }
  for (i = 0; i < np; i++)          /* SDV = dt . DV */
    for (j = 0; j < np; j++) {
      double s = 0.0;
      for (k = 0; k < np; k++) s += dt[i*np + k] * DV[k*np + j];
      SDV[i*np + j] = s;
    }
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) SDT[j*np + i] = SDV[i*np + j];

  free (z); free (w); free (DV); free (DT); free (FW); free (BW);
  free (S); free (dv); free (dt);
}

void orc_edgemaps (int np, int* emap, int* pmap)
/* Edge-major permutation, femlib/mapping.c:46-132: CCW around the boundary
 * from node 0, then the interior in row-major order; pmap = inverse.     */
{
  const int nm = np - 1;
  int i, j, k = 0, n = 0;
  emap[0] = 0;
  for (i = 1; i < np; i++) emap[++k] = n += 1;
  for (i = 1; i < np; i++) emap[++k] = n += np;
  for (i = 1; i < np; i++) emap[++k] = n -= 1;
  for (i = 1; i < np; i++) emap[++k] = n -= np;
  for (i = 1; i < nm; i++)
    for (j = 1; j < nm; j++) emap[k++] = i * np + j;
  for (i = 0; i < np * np; i++) pmap[emap[i]] = i;
}

/* ======================================================================== *
 * Mesh, geometry and numbering.
 * ======================================================================== */

struct orc_mesh {
  int     np, nel, next, nint, npnp, nPlane, planeSize;
  double *z, *w, *DV, *DT, *SDV, *SDT;
  int    *emap, *pmap;
  double *x, *y;
  double *Q1, *Q2, *Q3, *Q4, *Q8, *drdx, *dsdx, *drdy, *dsdy;
  int     cyl;              /* cylindrical: Q1-Q4 premultiplied by y, Q8 not */
  int    *hasQ3, *hasdrdx, *hasdsdx, *hasdrdy, *hasdsdy;
  int     nglobal, nsolve, nband;
  int    *btog, *bmask, *emask;
  int    *prov;             /* Mesh::buildMap numbering (before sortGid)   */
  int     nbound, *belmt, *bside, *bkind;
  double *barea, *imass;
};

#define RM(i, j, n) ((i) * (n) + (j))

static void mxm (const double* A, int nra, const double* B, int nca,
                 double* C, int ncb)
/* C = A . B, row-major (veclib/xmxm.f:32-60, dmxm).                       */
{
  int i, j, k;
  for (i = 0; i < nra; i++)
    for (j = 0; j < ncb; j++) {
      double s = 0.0;
      for (k = 0; k < nca; k++) s += A[i*nca + k] * B[k*ncb + j];
      C[i*ncb + j] = s;
    }
}

static void mxma (const double* A, int nra, const double* B, int nca,
                  double* C, int ncb)
/* C += A . B (veclib/xmxm.f, dmxma).                                      */
{
  int i, j, k;
  for (i = 0; i < nra; i++)
    for (j = 0; j < ncb; j++) {
      double s = 0.0;
      for (k = 0; k < nca; k++) s += A[i*nca + k] * B[k*ncb + j];
      C[i*ncb + j] += s;
    }
}

static void mesh_elmt (const orc_mesh* M, const double* vx, const double* vy,
                       double* x, double* y)
/* Straight-sided quad node generation, src/mesh.cpp:728-886 (meshSide,
 * meshElmt): sides by linear interpolation in the GLL spacing, sides 2-3
 * with reversed-negated spacing, interior by a Coons patch.              */
{
  const int np = M->np, nm = np - 1;
  const double* z = M->z;
  int i, j, s;
  double* work = xcalloc (np, sizeof (double));
  for (s = 0; s < 4; s++) {
    const double x1 = vx[s], y1 = vy[s], x2 = vx[(s+1)%4], y2 = vy[(s+1)%4];
    const double dx = x2 - x1, dy = y2 - y1;
    for (i = 0; i < np; i++) work[i] = (s < 2) ? z[i] : -z[nm - i];
    for (i = 0; i < nm; i++) {
      const double px = x1 + dx * 0.5 * (work[i] + 1.0);
      const double py = y1 + dy * 0.5 * (work[i] + 1.0);
      int k;
      switch (s) {
      case 0:  k = RM (0, i, np);       break;
      case 1:  k = RM (i, nm, np);      break;
      case 2:  k = RM (nm, nm - i, np); break;
      default: k = RM (nm - i, 0, np);  break;
      }
      x[k] = px; y[k] = py;
    }
  }
  for (i = 1; i < nm; i++)
    for (j = 1; j < nm; j++) {
      const double zr = z[j], zs = z[i];
      double* c[2] = { x, y };
      int q;
      for (q = 0; q < 2; q++) {
        double* t = c[q];
        t[RM (i, j, np)] =
          0.50 * ((1.0 - zr) * t[RM (i, 0, np)] + (1.0 + zr) * t[RM (i, nm, np)] +
                  (1.0 - zs) * t[RM (0, j, np)] + (1.0 + zs) * t[RM (nm, j, np)])
        - 0.25 * ((1.0 - zr) * (1.0 - zs) * t[RM (0, 0, np)] +
                  (1.0 - zr) * (1.0 + zs) * t[RM (nm, 0, np)] +
                  (1.0 - zs) * (1.0 + zr) * t[RM (0, nm, np)] +
                  (1.0 + zs) * (1.0 + zr) * t[RM (nm, nm, np)]);
      }
    }
  free (work);
}

static double nrm2 (int n, const double* x)
{
  double s = 0.0; int i;
  for (i = 0; i < n; i++) s += x[i] * x[i];
  return sqrt (s);
}

static void mesh_mapping (orc_mesh* M, int e)
/* Geometric factors, src/element.cpp:1281-1436 (Element::mapping), with
 * the cylindrical premultiplication of Q1-Q4 by y (:1417-1426; Q8 keeps
 * the unweighted mass) and the null-mapping flags of :1431-1435.         */
{
  const int np = M->np, npnp = M->npnp;
  const double EPS = 4 * EPSDP;
  const double *x = M->x + e*npnp, *y = M->y + e*npnp;
  double *dxdr = xcalloc (npnp, sizeof (double)), *dxds = xcalloc (npnp, sizeof (double));
  double *dydr = xcalloc (npnp, sizeof (double)), *dyds = xcalloc (npnp, sizeof (double));
  double *jac = xcalloc (npnp, sizeof (double));
  double *Q1 = M->Q1 + e*npnp, *Q2 = M->Q2 + e*npnp, *Q3 = M->Q3 + e*npnp;
  double *Q4 = M->Q4 + e*npnp, *drdx = M->drdx + e*npnp, *dsdx = M->dsdx + e*npnp;
  double *Q8 = M->Q8 + e*npnp;
  double *drdy = M->drdy + e*npnp, *dsdy = M->dsdy + e*npnp;
  int i, j;

  mxm (x, np, M->DT, np, dxdr, np);
  mxm (M->DV, np, x, np, dxds, np);
  mxm (y, np, M->DT, np, dydr, np);
  mxm (M->DV, np, y, np, dyds, np);
  for (i = 0; i < npnp; i++) {
    jac[i] = dxdr[i] * dyds[i] - dxds[i] * dydr[i];
    if (jac[i] < EPS) {
      fprintf (stderr, "orc mapping: Jacobian of element %d nonpositive\n", e + 1);
      abort ();
    }
  }
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) {
      const int k = RM (i, j, np);
      const double WW = M->w[i] * M->w[j];
      Q1[k] = (dxds[k] * dxds[k] + dyds[k] * dyds[k]) / jac[k] * WW;
      Q2[k] = (dxdr[k] * dxdr[k] + dydr[k] * dydr[k]) / jac[k] * WW;
      Q3[k] = (-dydr[k] * dyds[k] - dxdr[k] * dxds[k]) / jac[k] * WW;
      Q4[k] = jac[k] * WW;
      drdx[k] =  dyds[k] / jac[k];
      drdy[k] = -dxds[k] / jac[k];
      dsdx[k] = -dydr[k] / jac[k];
      dsdy[k] =  dxdr[k] / jac[k];
    }
  memcpy (Q8, Q4, npnp * sizeof (double));
  if (M->cyl)
    for (i = 0; i < npnp; i++) {
      Q1[i] *= y[i]; Q2[i] *= y[i]; Q3[i] *= y[i]; Q4[i] *= y[i];
    }
  M->hasdrdx[e] = nrm2 (npnp, drdx) >= EPS;
  M->hasdrdy[e] = nrm2 (npnp, drdy) >= EPS;
  M->hasdsdx[e] = nrm2 (npnp, dsdx) >= EPS;
  M->hasdsdy[e] = nrm2 (npnp, dsdy) >= EPS;
  M->hasQ3  [e] = nrm2 (npnp, Q3)   >= EPS;
  if (!M->hasdrdx[e]) memset (drdx, 0, npnp * sizeof (double));
  if (!M->hasdrdy[e]) memset (drdy, 0, npnp * sizeof (double));
  if (!M->hasdsdx[e]) memset (dsdx, 0, npnp * sizeof (double));
  if (!M->hasdsdy[e]) memset (dsdy, 0, npnp * sizeof (double));
  if (!M->hasQ3  [e]) memset (Q3,   0, npnp * sizeof (double));
  free (dxdr); free (dxds); free (dydr); free (dyds); free (jac);
}

static int element_band (const orc_mesh* M, const int* btog, const int* bmask,
                         int nel, int next)
/* utility/enumerate.cpp:903-943 (globalBandwidth / bandwidthSC).          */
{
  int e, i, nband = 0;
  for (e = 0; e < nel; e++) {
    int lo = 1 << 30, hi = -1;
    for (i = 0; i < next; i++)
      if (!bmask[e*next + i]) {
        const int g = btog[e*next + i];
        if (g < lo) lo = g;
        if (g > hi) hi = g;
      }
    if (hi >= 0 && hi - lo > nband) nband = hi - lo;
  }
  (void) M;
  return nband + 1;
}

static int uf_find (int* par, int a)
{
  while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; }
  return a;
}

static void mesh_number (orc_mesh* M, const int* nbr, const int* skind)
/* Global numbering of element-boundary nodes.  Provisional numbers by an
 * element/side traverse (src/mesh.cpp:889-972, Mesh::buildMap: joined sides
 * share their knots, traversed in opposite directions; corners shared by
 * transitivity), essential nodes moved last keeping relative order
 * (utility/enumerate.cpp:621-677, sortGid), then a Cuthill-McKee
 * relabelling of the unknowns, kept only if it narrows the band (the role
 * of utility/enumerate.cpp:680-809).                                    */
{
  const int np = M->np, nel = M->nel, next = M->next, ni = np - 2, nm = np - 1;
  const int nside = 4 * nel;
  int e, s, i, k, ng = 0;
  int *par, *cid, *sfirst, *prov, *ess, *perm;

  par = xcalloc (nside, sizeof (int));
  cid = xcalloc (nside, sizeof (int));
  sfirst = xcalloc (nside, sizeof (int));
  for (i = 0; i < nside; i++) { par[i] = i; cid[i] = -1; }
  for (i = 0; i < nside; i++) {
    const int j = nbr[i];
    if (j > i) {
      const int e1 = i / 4, s1 = i % 4, e2 = j / 4, s2 = j % 4;
      int ra = uf_find (par, 4*e1 + s1), rb = uf_find (par, 4*e2 + (s2 + 1) % 4);
      par[ra] = rb;
      ra = uf_find (par, 4*e1 + (s1 + 1) % 4); rb = uf_find (par, 4*e2 + s2);
      par[ra] = rb;
    }
  }
  prov = xcalloc (nel * next, sizeof (int));
  for (e = 0; e < nel; e++)
    for (s = 0; s < 4; s++) {
      const int i0 = 4*e + s, r = uf_find (par, i0), j = nbr[i0];
      if (cid[r] < 0) cid[r] = ng++;
      prov[e*next + s*nm] = cid[r];
      if (j >= 0 && j < i0) {
        for (k = 0; k < ni; k++) prov[e*next + s*nm + 1 + k] = sfirst[j] + ni - 1 - k;
        sfirst[i0] = sfirst[j];
      } else {
        sfirst[i0] = ng;
        for (k = 0; k < ni; k++) prov[e*next + s*nm + 1 + k] = ng++;
      }
    }
  free (par); free (cid); free (sfirst);
  M->prov = xcalloc (nel * next, sizeof (int));
  memcpy (M->prov, prov, nel * next * sizeof (int));

  /* -- Essential mask on global nodes.                                    */
  ess = xcalloc (ng, sizeof (int));
  M->nbound = 0;
  for (e = 0; e < nel; e++)
    for (s = 0; s < 4; s++) {
      if (skind[4*e + s]) M->nbound++;
      if (skind[4*e + s] == 1)
        for (i = 0; i <= nm; i++)
          ess[prov[e*next + ((s*nm + i) % next)]] = 1;
    }

  /* -- sortGid: unknowns first (ascending), then essential (ascending).  */
  perm = xcalloc (ng, sizeof (int));
  {
    int nu = 0, ne = 0, nunk = 0;
    for (i = 0; i < ng; i++) if (!ess[i]) nunk++;
    for (i = 0; i < ng; i++) perm[i] = ess[i] ? nunk + ne++ : nu++;
    M->nsolve = nunk;
  }
  M->nglobal = ng;
  M->btog  = xcalloc (nel * next, sizeof (int));
  M->bmask = xcalloc (nel * next, sizeof (int));
  M->emask = xcalloc (nel, sizeof (int));
  for (i = 0; i < nel * next; i++) {
    M->btog [i] = perm[prov[i]];
    M->bmask[i] = ess[prov[i]];
  }
  for (e = 0; e < nel; e++)
    for (i = 0; i < next; i++) if (M->bmask[e*next + i]) M->emask[e] = 1;
  M->nband = element_band (M, M->btog, M->bmask, nel, next);

  /* -- Cuthill-McKee on the unknown-node adjacency (own simple variant:
   *    BFS from the lowest-degree node, neighbours by increasing degree,
   *    reversed).                                                         */
  if (M->nsolve > 2) {
    const int n = M->nsolve;
    int *deg = xcalloc (n, sizeof (int)), *ptr = xcalloc (n + 1, sizeof (int));
    int *adj, *fill, *order, *seen, *newid, *trial, head, tail, nb2;
    int cap = 0;
    for (e = 0; e < nel; e++) {
      int cnt = 0;
      for (i = 0; i < next; i++) if (M->btog[e*next + i] < n) cnt++;
      for (i = 0; i < next; i++) {
        const int g = M->btog[e*next + i];
        if (g < n) deg[g] += cnt - 1;
      }
    }
    for (i = 0; i < n; i++) { ptr[i+1] = ptr[i] + deg[i]; cap += deg[i]; }
    adj  = xcalloc (cap, sizeof (int));
    fill = xcalloc (n, sizeof (int));
    for (e = 0; e < nel; e++)
      for (i = 0; i < next; i++) {
        const int g = M->btog[e*next + i];
        if (g >= n) continue;
        for (k = 0; k < next; k++) {
          const int h = M->btog[e*next + k];
          if (h < n && h != g) adj[ptr[g] + fill[g]++] = h;
        }
      }
    /* degree counted with duplicates; that only biases the ordering.     */
    order = xcalloc (n, sizeof (int));
    seen  = xcalloc (n, sizeof (int));
    newid = xcalloc (n, sizeof (int));
    trial = xcalloc (nel * next, sizeof (int));
    head = tail = 0;
    while (tail < n) {
      int root = -1;
      for (i = 0; i < n; i++)
        if (!seen[i] && (root < 0 || deg[i] < deg[root])) root = i;
      seen[root] = 1; order[tail++] = root;
      while (head < tail) {
        const int g = order[head++];
        const int t0 = tail;
        for (k = ptr[g]; k < ptr[g] + fill[g]; k++) {
          const int h = adj[k];
          if (!seen[h]) { seen[h] = 1; order[tail++] = h; }
        }
        /* sort the newly added by degree (insertion sort)                */
        for (i = t0 + 1; i < tail; i++) {
          const int v = order[i]; int j2 = i - 1;
          while (j2 >= t0 && deg[order[j2]] > deg[v]) { order[j2+1] = order[j2]; j2--; }
          order[j2+1] = v;
        }
      }
    }
    for (i = 0; i < n; i++) newid[order[i]] = n - 1 - i;
    for (i = 0; i < nel * next; i++) {
      const int g = M->btog[i];
      trial[i] = (g < n) ? newid[g] : g;
    }
    nb2 = element_band (M, trial, M->bmask, nel, next);
    if (nb2 < M->nband) {
      memcpy (M->btog, trial, nel * next * sizeof (int));
      M->nband = nb2;
    }
    free (deg); free (ptr); free (adj); free (fill); free (order);
    free (seen); free (newid); free (trial);
  }

  /* -- Boundary list in (element, side) order.                          */
  M->belmt = xcalloc (M->nbound, sizeof (int));
  M->bside = xcalloc (M->nbound, sizeof (int));
  M->bkind = xcalloc (M->nbound, sizeof (int));
  k = 0;
  for (e = 0; e < nel; e++)
    for (s = 0; s < 4; s++)
      if (skind[4*e + s]) {
        M->belmt[k] = e; M->bside[k] = s; M->bkind[k] = skind[4*e + s]; k++;
      }

  free (prov); free (ess); free (perm);
}

static int side_node (int np, int s, int q);

static void mesh_weights (orc_mesh* M)
/* Inverse global mass (unweighted Q8) and side quadrature weights.       */
{
  const int np = M->np, npnp = M->npnp, nel = M->nel;
  int e, b, i;
  free (M->imass); free (M->barea);
  /* -- Inverse global mass (src/bcmgr.cpp:806-822).                      */
  M->imass = xcalloc (M->nglobal, sizeof (double));
  for (e = 0; e < nel; e++)
    for (i = 0; i < M->next; i++)
      M->imass[M->btog[e*M->next + i]] += M->Q8[e*npnp + M->emap[i]];
  for (i = 0; i < M->nglobal; i++) M->imass[i] = 1.0 / M->imass[i];

  /* -- Side quadrature weights (src/element.cpp:1653-1744, sideGeom).   */
  M->barea = xcalloc (M->nbound * np, sizeof (double));
  for (b = 0; b < M->nbound; b++) {
    const int s = M->bside[b], el = M->belmt[b];
    const double *x = M->x + el*npnp, *y = M->y + el*npnp;
    for (i = 0; i < np; i++) {
      double tx = 0.0, ty = 0.0; int k;
      for (k = 0; k < np; k++) {
        int idx;
        switch (s) {
        case 0:  idx = RM (0, k, np);  break;
        case 1:  idx = RM (k, np-1, np); break;
        case 2:  idx = RM (np-1, k, np); break;
        default: idx = RM (k, 0, np);  break;
        }
        tx += M->DV[i*np + k] * x[idx];
        ty += M->DV[i*np + k] * y[idx];
      }
      M->barea[b*np + i] = sqrt (tx*tx + ty*ty) * M->w[i];
      /* cylindrical: times y at the side's i-th node, sides run CCW
         (src/edge.cpp:75 with Element::sideGeom's node order)          */
      if (M->cyl) M->barea[b*np + i] *= y[side_node (np, s, i)];
    }
  }
}

int orc_mesh_set_cylindrical (orc_mesh* M)
/* The CYLINDRICAL token: factors and side weights from y (x axial, y
 * radial, no node at y < 0).                                            */
{
  int e, i;
  for (i = 0; i < M->nPlane; i++) if (M->y[i] < 0.0) return 1;
  M->cyl = 1;
  for (e = 0; e < M->nel; e++) mesh_mapping (M, e);
  mesh_weights (M);
  return 0;
}

orc_mesh* orc_mesh_create (int np, int nel, const double* vx, const double* vy,
                           const int* nbr, const int* skind,
                           int svv_mn, double svv_eps, double kinvis)
{
  orc_mesh* M = xcalloc (1, sizeof (orc_mesh));
  const int npnp = np * np;
  int e;

  M->np = np; M->nel = nel; M->npnp = npnp;
  M->next = 4 * (np - 1); M->nint = npnp - M->next;
  M->nPlane = nel * npnp;
  M->planeSize = M->nPlane + (M->nPlane % 2);      /* src/geometry.cpp:107 */

  M->z = xcalloc (np, sizeof (double)); M->w = xcalloc (np, sizeof (double));
  M->DV = xcalloc (npnp, sizeof (double)); M->DT = xcalloc (npnp, sizeof (double));
  M->SDV = xcalloc (npnp, sizeof (double)); M->SDT = xcalloc (npnp, sizeof (double));
  M->emap = xcalloc (npnp, sizeof (int)); M->pmap = xcalloc (npnp, sizeof (int));
  orc_zwglj (M->z, M->w, np);
  orc_dglj (M->DV, M->DT, M->z, np);
  orc_svv_operators (np, svv_mn, svv_eps, kinvis, M->SDV, M->SDT);
  orc_edgemaps (np, M->emap, M->pmap);

  M->x = xcalloc (nel * npnp, sizeof (double));
  M->y = xcalloc (nel * npnp, sizeof (double));
  M->Q1 = xcalloc (nel * npnp, sizeof (double)); M->Q2 = xcalloc (nel * npnp, sizeof (double));
  M->Q3 = xcalloc (nel * npnp, sizeof (double)); M->Q4 = xcalloc (nel * npnp, sizeof (double));
  M->Q8 = xcalloc (nel * npnp, sizeof (double));
  M->drdx = xcalloc (nel * npnp, sizeof (double)); M->dsdx = xcalloc (nel * npnp, sizeof (double));
  M->drdy = xcalloc (nel * npnp, sizeof (double)); M->dsdy = xcalloc (nel * npnp, sizeof (double));
  M->hasQ3 = xcalloc (nel, sizeof (int)); M->hasdrdx = xcalloc (nel, sizeof (int));
  M->hasdsdx = xcalloc (nel, sizeof (int)); M->hasdrdy = xcalloc (nel, sizeof (int));
  M->hasdsdy = xcalloc (nel, sizeof (int));

  for (e = 0; e < nel; e++) {
    mesh_elmt (M, vx + 4*e, vy + 4*e, M->x + e*npnp, M->y + e*npnp);
    mesh_mapping (M, e);
  }
  mesh_number (M, nbr, skind);

  mesh_weights (M);
  return M;
}

void orc_mesh_destroy (orc_mesh* M)
{
  if (!M) return;
  free (M->z); free (M->w); free (M->DV); free (M->DT); free (M->SDV); free (M->SDT);
  free (M->emap); free (M->pmap); free (M->x); free (M->y);
  free (M->Q1); free (M->Q2); free (M->Q3); free (M->Q4); free (M->Q8);
  free (M->drdx); free (M->dsdx); free (M->drdy); free (M->dsdy);
  free (M->hasQ3); free (M->hasdrdx); free (M->hasdsdx); free (M->hasdrdy); free (M->hasdsdy);
  free (M->btog); free (M->bmask); free (M->emask); free (M->prov);
  free (M->belmt); free (M->bside); free (M->bkind); free (M->barea); free (M->imass);
  free (M);
}

void orc_mesh_info (const orc_mesh* M, int* info)
{
  info[0] = M->np; info[1] = M->nel; info[2] = M->next; info[3] = M->nint;
  info[4] = M->nglobal; info[5] = M->nsolve; info[6] = M->nband;
  info[7] = M->nbound; info[8] = M->nPlane; info[9] = M->planeSize;
}

void orc_mesh_nodes (const orc_mesh* M, double* x, double* y)
{
  memcpy (x, M->x, M->nel * M->npnp * sizeof (double));
  memcpy (y, M->y, M->nel * M->npnp * sizeof (double));
}

void orc_mesh_bsides (const orc_mesh* M, int* belmt, int* bside)
{
  memcpy (belmt, M->belmt, M->nbound * sizeof (int));
  memcpy (bside, M->bside, M->nbound * sizeof (int));
}

int orc_mesh_renumber (orc_mesh* M, const int* btog, const int* bmask, int nglobal,
                       int nsolve)
/* Replace the numbering by a given one (session.num as BCmgr::buildnum
 * reads it, src/bcmgr.cpp:586-824): same partition of element-boundary
 * slots into global nodes, same essential mask, unknowns below nsolve.
 * Recomputes emask, the band (utility/enumerate.cpp:903-943) and the
 * inverse global mass (src/bcmgr.cpp:806-822).  Returns 0, or 1 for a
 * numbering that does not fit the mesh (then nothing is changed).      */
{
  const int nel = M->nel, next = M->next, ns = nel * next;
  int *to, *from, i, e, ok = 1;
  if (nglobal != M->nglobal || nsolve != M->nsolve) return 1;
  to = xcalloc (nglobal, sizeof (int));
  from = xcalloc (nglobal, sizeof (int));
  for (i = 0; i < nglobal; i++) to[i] = from[i] = -1;
  for (i = 0; i < ns && ok; i++) {
    const int g = btog[i], p = M->btog[i];
    if (g < 0 || g >= nglobal || bmask[i] != M->bmask[i] || (bmask[i] != 0) != (g >= nsolve))
      ok = 0;
    else if ((to[p] >= 0 && to[p] != g) || (from[g] >= 0 && from[g] != p))
      ok = 0;
    else { to[p] = g; from[g] = p; }
  }
  free (to); free (from);
  if (!ok) return 1;
  memcpy (M->btog, btog, ns * sizeof (int));
  for (e = 0; e < nel; e++) {
    M->emask[e] = 0;
    for (i = 0; i < next; i++) if (M->bmask[e*next + i]) M->emask[e] = 1;
  }
  M->nband = element_band (M, M->btog, M->bmask, nel, next);
  for (i = 0; i < nglobal; i++) M->imass[i] = 0.0;
  for (e = 0; e < nel; e++)
    for (i = 0; i < next; i++)
      M->imass[M->btog[e*next + i]] += M->Q8[e*M->npnp + M->emap[i]];
  for (i = 0; i < nglobal; i++) M->imass[i] = 1.0 / M->imass[i];
  return 0;
}

/* -- enumerate's numbering ------------------------------------------------
 * The node-numbering utility semtex runs for session.num (utility/
 * enumerate.cpp:540-943): the buildMap traverse numbering (src/mesh.cpp:
 * 889-1015, the provisional numbers here), unknowns moved before the
 * essential nodes keeping their order (Nsys::sortGid), then for
 * optimisation level 1 the SPARSPAK pseudo-peripheral root search from
 * node 1 and reverse Cuthill-McKee (George & Liu 1981; femlib/netlib.f:
 * 1543-1900) on the adjacency lists of Nsys::connectivSC.               */

static void sp_rootls (int root, const int* xadj, const int* adj, int* mask, int* nlvl,
                       int* xls, int* ls)
/* ROOTLS: level structure rooted at root over nodes with mask != 0 (all
 * arrays 1-based as in the reference)                                   */
{
  int lbegin, lvlend = 0, ccsize = 1, i, j;
  mask[root] = 0;
  ls[1] = root;
  *nlvl = 0;
  for (;;) {
    lbegin = lvlend + 1;
    lvlend = ccsize;
    *nlvl += 1;
    xls[*nlvl] = lbegin;
    for (i = lbegin; i <= lvlend; i++) {
      const int node = ls[i];
      for (j = xadj[node]; j <= xadj[node + 1] - 1; j++) {
        const int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        ccsize++;
        ls[ccsize] = nbr;
        mask[nbr] = 0;
      }
    }
    if (ccsize - lvlend <= 0) break;
  }
  xls[*nlvl + 1] = lvlend + 1;
  for (i = 1; i <= ccsize; i++) mask[ls[i]] = 1;
}

static void sp_fnroot (int* root, const int* xadj, const int* adj, int* mask, int* nlvl,
                       int* xls, int* ls)
/* FNROOT: pseudo-peripheral node of the component of root (modified
 * Gibbs-Poole-Stockmeyer)                                               */
{
  int ccsize, nunlvl, j, k;
  sp_rootls (*root, xadj, adj, mask, nlvl, xls, ls);
  ccsize = xls[*nlvl + 1] - 1;
  if (*nlvl == 1 || *nlvl == ccsize) return;
  for (;;) {
    const int jstrt = xls[*nlvl];
    int mindeg = ccsize;
    *root = ls[jstrt];
    if (ccsize != jstrt)
      for (j = jstrt; j <= ccsize; j++) {
        const int node = ls[j];
        int ndeg = 0;
        for (k = xadj[node]; k <= xadj[node + 1] - 1; k++)
          if (mask[adj[k]] > 0) ndeg++;
        if (ndeg < mindeg) { *root = node; mindeg = ndeg; }
      }
    sp_rootls (*root, xadj, adj, mask, &nunlvl, xls, ls);
    if (nunlvl <= *nlvl) return;
    *nlvl = nunlvl;
    if (*nlvl >= ccsize) return;
  }
}

static void sp_degree (int root, int* xadj, const int* adj, const int* mask, int* deg,
                       int* ccsize, int* ls)
/* DEGREE: degrees within the masked component of root (xadj signs mark
 * visited nodes and are restored)                                       */
{
  int lbegin, lvlend = 0, i, j;
  ls[1] = root;
  xadj[root] = -xadj[root];
  *ccsize = 1;
  for (;;) {
    lbegin = lvlend + 1;
    lvlend = *ccsize;
    for (i = lbegin; i <= lvlend; i++) {
      const int node = ls[i];
      const int jstrt = -xadj[node], jstop = abs (xadj[node + 1]) - 1;
      int ideg = 0;
      for (j = jstrt; j <= jstop; j++) {
        const int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        ideg++;
        if (xadj[nbr] < 0) continue;
        xadj[nbr] = -xadj[nbr];
        (*ccsize)++;
        ls[*ccsize] = nbr;
      }
      deg[node] = ideg;
    }
    if (*ccsize - lvlend <= 0) break;
  }
  for (i = 1; i <= *ccsize; i++) xadj[ls[i]] = -xadj[ls[i]];
}

static void sp_rcm (int root, int* xadj, const int* adj, int* mask, int* perm, int* ccsize,
                    int* deg)
/* RCM: reverse Cuthill-McKee numbering of the component of root from root,
 * neighbours of each node taken by increasing degree (linear insertion)  */
{
  int lbegin, lvlend = 0, lnbr = 1, i, j, k, l;
  sp_degree (root, xadj, adj, mask, deg, ccsize, perm);
  mask[root] = 0;
  if (*ccsize <= 1) return;
  do {
    lbegin = lvlend + 1;
    lvlend = lnbr;
    for (i = lbegin; i <= lvlend; i++) {
      const int node = perm[i];
      const int fnbr = lnbr + 1;
      for (j = xadj[node]; j <= xadj[node + 1] - 1; j++) {
        const int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        lnbr++;
        mask[nbr] = 0;
        perm[lnbr] = nbr;
      }
      if (fnbr >= lnbr) continue;
      k = fnbr;
      do {
        int nbr;
        l = k;
        k++;
        nbr = perm[k];
        while (l >= fnbr && deg[perm[l]] > deg[nbr]) { perm[l + 1] = perm[l]; l--; }
        perm[l + 1] = nbr;
      } while (k < lnbr);
    }
  } while (lnbr > lvlend);
  for (i = 1, l = *ccsize; i <= *ccsize / 2; i++, l--) {
    const int t = perm[l]; perm[l] = perm[i]; perm[i] = t;
  }
}

int orc_mesh_enumerate (orc_mesh* M, int optlevel)
/* Renumber M as `enumerate -O<optlevel>` numbers it (levels 0-3; the
 * utility's default, and the regression sessions', is 3).
 * Returns 0, or 1 for an unsupported level.                             */
{
  const int nel = M->nel, next = M->next, ns = nel * next, ng = M->nglobal;
  int *ess, *newid, *bt, i, e, k;
  int nunk = 0, ne = 0;
  if (optlevel < 0 || optlevel > 3) return 1;
  ess = xcalloc (ng, sizeof (int));
  newid = xcalloc (ng, sizeof (int));
  bt = xcalloc (ns, sizeof (int));
  for (i = 0; i < ns; i++) ess[M->prov[i]] = M->bmask[i];
  for (i = 0; i < ng; i++) nunk += !ess[i];
  for (i = 0, k = 0; i < ng; i++) newid[i] = ess[i] ? nunk + ne++ : k++;   /* sortGid */
  for (i = 0; i < ns; i++) bt[i] = newid[M->prov[i]];
  if (optlevel >= 1 && nunk > 0) {
    /* adjacency lists (connectivSC): per element, for each unknown slot i,
     * the other unknown slots' nodes in slot order, first appearance only */
    int* cnt = xcalloc (nunk, sizeof (int));
    int* cap = xcalloc (nunk, sizeof (int));
    int** lst = xcalloc (nunk, sizeof (int*));
    int *xadj, *adj, *mask, *xls, *perm, *invperm, tab = 0, root = 1, nlvl, ccs;
    for (e = 0; e < nel; e++) {
      const int* bm = bt + e * next;
      const int* mk = M->bmask + e * next;
      for (i = 0; i < next; i++) {
        int j;
        if (mk[i]) continue;
        for (j = 0; j < next; j++) {
          const int g = bm[i], h = bm[j];
          int q, found = 0;
          if (i == j || mk[j]) continue;
          for (q = 0; q < cnt[g] && !found; q++) found = lst[g][q] == h;
          if (found) continue;
          if (cnt[g] == cap[g]) {
            cap[g] = cap[g] ? 2 * cap[g] : 16;
            lst[g] = realloc (lst[g], cap[g] * sizeof (int));
          }
          lst[g][cnt[g]++] = h;
        }
      }
    }
    for (i = 0; i < nunk; i++) tab += cnt[i];
    xadj = xcalloc (nunk + 2, sizeof (int));
    adj = xcalloc (tab + 2, sizeof (int));
    mask = xcalloc (nunk + 2, sizeof (int));
    xls = xcalloc (nunk + 2, sizeof (int));
    perm = xcalloc (nunk + 2, sizeof (int));
    invperm = xcalloc (ng, sizeof (int));
    for (i = 0, k = 1; i < nunk; i++) {            /* fillAdjncy, 1-based */
      int q;
      xadj[i + 1] = k;
      for (q = 0; q < cnt[i]; q++, k++) adj[k] = lst[i][q] + 1;
      free (lst[i]);
    }
    xadj[nunk + 1] = k;
    if (optlevel == 1) {
      for (i = 1; i <= nunk; i++) mask[i] = 1;
      sp_fnroot (&root, xadj, adj, mask, &nlvl, xls, perm);
    } else {
      /* level 2: fnroot from trial roots 1, 11, 21, ...; level 3: every
       * unknown as the RCM root.  Keep the first root of least global
       * bandwidth (Nsys::globalBandwidth: max over elements of the
       * unknowns' index span, plus one).                                 */
      int best = 1, bwmin = INT_MAX, r;
      for (r = 1; r <= nunk; r += (optlevel == 2) ? 10 : 1) {
        int rt = r, bw = 0;
        for (i = 1; i <= nunk; i++) mask[i] = 1;
        if (optlevel == 2) {
          sp_fnroot (&rt, xadj, adj, mask, &nlvl, xls, perm);
          for (i = 1; i <= nunk; i++) mask[i] = 1;  /* as left by fnroot */
        }
        sp_rcm (rt, xadj, adj, mask, perm, &ccs, xls);
        for (i = 0; i < nunk; i++) invperm[perm[i + 1] - 1] = i;
        for (e = 0; e < nel; e++) {
          int lo = INT_MAX, hi = INT_MIN;
          for (k = 0; k < next; k++) {
            int g;
            if (M->bmask[e*next + k]) continue;
            g = invperm[bt[e*next + k]];
            lo = g < lo ? g : lo; hi = g > hi ? g : hi;
          }
          if (hi >= lo && hi - lo > bw) bw = hi - lo;
        }
        if (bw + 1 < bwmin) { bwmin = bw + 1; best = rt; }
      }
      root = best;
    }
    for (i = 1; i <= nunk; i++) mask[i] = 1;
    sp_rcm (root, xadj, adj, mask, perm, &ccs, xls);
    for (i = nunk; i < ng; i++) invperm[i] = i;
    for (i = 0; i < nunk; i++) invperm[perm[i + 1] - 1] = i;
    for (i = 0; i < ns; i++) bt[i] = invperm[bt[i]];
    free (cnt); free (cap); free (lst); free (xadj); free (adj); free (mask);
    free (xls); free (perm); free (invperm);
  }
  memcpy (M->btog, bt, ns * sizeof (int));
  free (ess); free (newid); free (bt);
  for (e = 0; e < nel; e++) {
    M->emask[e] = 0;
    for (i = 0; i < next; i++) if (M->bmask[e*next + i]) M->emask[e] = 1;
  }
  M->nband = element_band (M, M->btog, M->bmask, nel, next);
  for (i = 0; i < ng; i++) M->imass[i] = 0.0;
  for (e = 0; e < nel; e++)
    for (i = 0; i < next; i++)
      M->imass[M->btog[e*next + i]] += M->Q8[e*M->npnp + M->emap[i]];
  for (i = 0; i < ng; i++) M->imass[i] = 1.0 / M->imass[i];
  return 0;
}

void orc_mesh_numbering (const orc_mesh* M, int* btog, int* bmask, int* emask)
{
  memcpy (btog, M->btog, M->nel * M->next * sizeof (int));
  memcpy (bmask, M->bmask, M->nel * M->next * sizeof (int));
  memcpy (emask, M->emask, M->nel * sizeof (int));
}

void orc_mesh_factors (const orc_mesh* M, double* Q1, double* Q2, double* Q3,
                       double* Q4, double* drdx, double* dsdx, double* drdy,
                       double* dsdy, int* flags)
{
  const size_t n = (size_t) M->nel * M->npnp * sizeof (double);
  int e;
  memcpy (Q1, M->Q1, n); memcpy (Q2, M->Q2, n); memcpy (Q3, M->Q3, n);
  memcpy (Q4, M->Q4, n); memcpy (drdx, M->drdx, n); memcpy (dsdx, M->dsdx, n);
  memcpy (drdy, M->drdy, n); memcpy (dsdy, M->dsdy, n);
  for (e = 0; e < M->nel; e++) {
    flags[5*e + 0] = M->hasdrdx[e]; flags[5*e + 1] = M->hasdsdx[e];
    flags[5*e + 2] = M->hasdrdy[e]; flags[5*e + 3] = M->hasdsdy[e];
    flags[5*e + 4] = M->hasQ3[e];
  }
}

void orc_mesh_operators (const orc_mesh* M, double* z, double* w, double* DV,
                         double* SDV, int* emap)
{
  memcpy (z, M->z, M->np * sizeof (double));
  memcpy (w, M->w, M->np * sizeof (double));
  memcpy (DV, M->DV, M->npnp * sizeof (double));
  memcpy (SDV, M->SDV, M->npnp * sizeof (double));
  memcpy (emap, M->emap, M->npnp * sizeof (int));
}

void orc_mesh_imass (const orc_mesh* M, double* imass)
{ memcpy (imass, M->imass, M->nglobal * sizeof (double)); }

void orc_mesh_barea (const orc_mesh* M, double* barea)
{ memcpy (barea, M->barea, M->nbound * M->np * sizeof (double)); }

/* ======================================================================== *
 * Element operators.
 * ======================================================================== */

/* -- Fourier-direction SVV weights (SVV::coeffs_z, src/svv.cpp:82-123):
 *    S[k - base] = 1 + eps/nu exp(-((N-k)/(mz-k))^2) for modes k > mz of
 *    base..base+nummodes-1, N = nmode - 1; 1 elsewhere, and all 1 unless
 *    0 <= mz < N and eps > EPSDP.  The solve settings below are the tokens
 *    SVV_MZ, SVV_EPSZ, KINVIS and Geometry::nMode (0: base + nummodes).   */
static int    svvz_mz = -1, svvz_nmode = 0;
static double svvz_eps = 0.0, svvz_nu = 1.0;

void orc_svv_coeffs_z (int nummodes, int nmode, int base, int mz, double eps, double nu,
                       double* S)
{
  const int N = nmode - 1;
  int i;
  for (i = 0; i < nummodes; i++) S[i] = 0.0;
  if (mz >= 0 && mz < N && eps > EPSDP) {
    for (i = base; i < base + nummodes; i++)
      if (i > mz) {
        const double q = (N - i) / (1.0 * mz - i);
        S[i - base] = exp (-q * q);
      }
    for (i = 0; i < nummodes; i++) S[i] = 1.0 + eps / nu * S[i];
  } else
    for (i = 0; i < nummodes; i++) S[i] = 1.0;
}

void orc_set_svv_z (int mz, double eps, double nu, int nmode)
{
  svvz_mz = mz; svvz_eps = eps; svvz_nu = nu; svvz_nmode = nmode;
}

static char field_name = 'c';

void orc_set_field_name (int name) { field_name = (char) name; }

static double mode_constant (const orc_mesh* M, int mode, double beta)
/* Field::modeConstant (src/field.cpp:1121-1146) for the field set by
 * orc_set_field_name ('c' unless set).                                  */
{
  if (!M->cyl || field_name == 'c' || field_name == 'p' || field_name == 'u')
    return beta * mode;
  if (field_name == 'v') return (mode == 0) ? 1.0 : beta * mode + 1.0;
  return (mode == 0) ? 1.0 : beta * mode - 1.0;
}

void orc_couple (double* v, double* w, int nz, int n, int base, int dir)
/* AuxField::couple / Field::coupleBCs (src/auxfield.cpp:1352-1420,
 * src/field.cpp:1043-1118), the Veclib sequence restated: per mode k from
 * kLo (1 when the planes start at mode 0), planes 2k and 2k+1.           */
{
  const int nmode = (nz > 1) ? nz / 2 : 0, klo = (base == 0) ? 1 : 0;
  int k, i;
  double* tp = xcalloc (n, sizeof (double));
  for (k = klo; k < nmode; k++) {
    double *Vr = v + (size_t)(2*k) * n, *Vi = Vr + n;
    double *Wr = w + (size_t)(2*k) * n, *Wi = Wr + n;
    if (dir > 0) {
      for (i = 0; i < n; i++) tp[i] = Vr[i];
      for (i = 0; i < n; i++) Vr[i] = Vr[i] - Wi[i];        /* vsub        */
      for (i = 0; i < n; i++) Wi[i] = Wi[i] + tp[i];        /* vadd        */
      for (i = 0; i < n; i++) tp[i] = Wr[i];
      for (i = 0; i < n; i++) Wr[i] = Wi[i];
      for (i = 0; i < n; i++) Wi[i] = Vi[i] - tp[i];        /* vsub        */
      for (i = 0; i < n; i++) Vi[i] = Vi[i] + tp[i];        /* vadd        */
    } else {
      for (i = 0; i < n; i++) tp[i] = Vr[i];
      for (i = 0; i < n; i++) Vr[i] = 0.5 * (Vr[i] + Wr[i]); /* svvpt      */
      for (i = 0; i < n; i++) Wr[i] = 0.5 * (Wr[i] - tp[i]); /* svvmt      */
      for (i = 0; i < n; i++) tp[i] = Wi[i];
      for (i = 0; i < n; i++) Wi[i] = Wr[i];
      for (i = 0; i < n; i++) Wr[i] = 0.5 * (Vi[i] - tp[i]); /* svvmt      */
      for (i = 0; i < n; i++) Vi[i] = 0.5 * (Vi[i] + tp[i]); /* svvpt      */
    }
  }
  free (tp);
}

static void pick_ops (const orc_mesh* M, double lambda2, const double** dv,
                      const double** dt)
/* SVV operators whenever lambda2 > EPSDP (src/element.cpp:1484,1532,1640). */
{
  if (lambda2 > EPSDP) { *dv = M->SDV; *dt = M->SDT; }
  else                 { *dv = M->DV;  *dt = M->DT;  }
}

void orc_helm_op (const orc_mesh* M, int e, double lambda2, double betak2,
                  const double* src, double* tgt)
/* src/element.cpp:1621-1650 (HelmholtzOp) with the kernel of :1564-1618
 * (HelmholtzKern; cylindrical: hCon = betak2/r^2 + lambda2 per node, 0
 * where r^2 <= EPSDP).  src and tgt may alias.                           */
{
  const int np = M->np, npnp = M->npnp;
  const double *dv, *dt;
  const double *g1 = M->Q1 + e*npnp, *g2 = M->Q2 + e*npnp, *g3 = M->Q3 + e*npnp;
  const double *g4 = M->Q4 + e*npnp, *r = M->y + e*npnp;
  double R[32*32], S[32*32], T[32*32], hc[32*32];
  int ij;

  for (ij = 0; ij < npnp; ij++) {
    const double r2 = r[ij] * r[ij];
    hc[ij] = M->cyl ? ((r2 > EPSDP) ? (betak2 / r2 + lambda2) : 0.0) : betak2 + lambda2;
  }

  pick_ops (M, lambda2, &dv, &dt);
  mxm (src, np, dt, np, R, np);
  mxm (dv, np, src, np, S, np);
  if (M->hasQ3[e])
    for (ij = 0; ij < npnp; ij++) {
      const double tmp = R[ij];
      R[ij] = g1[ij] * R[ij] + g3[ij] * S[ij];
      S[ij] = g2[ij] * S[ij] + g3[ij] * tmp;
      T[ij] = g4[ij] * src[ij] * hc[ij];
    }
  else
    for (ij = 0; ij < npnp; ij++) {
      R[ij] *= g1[ij];
      S[ij] *= g2[ij];
      T[ij] = g4[ij] * src[ij] * hc[ij];
    }
  mxma (dt, np, S, np, T, np);
  mxma (R, np, dv, np, T, np);
  memcpy (tgt, T, npnp * sizeof (double));
}

static void helm_row (const orc_mesh* M, int e, double lambda2, double betak2,
                      int i, int j, double* hij)
/* src/element.cpp:1439-1509 (HelmholtzRow; cylindrical: hCon = betak2/r^2
 * + lambda2 at node (i, j) where r^2 > EPSDP).                           */
{
  const int np = M->np, npnp = M->npnp;
  const double *dvr, *dtr;
  const double *Q1 = M->Q1 + e*npnp, *Q2 = M->Q2 + e*npnp, *Q3 = M->Q3 + e*npnp;
  const double *Q4 = M->Q4 + e*npnp;
  const double r2 = M->y[e*npnp + RM (i, j, np)] * M->y[e*npnp + RM (i, j, np)];
  const double hCon = (M->cyl && r2 > EPSDP) ? (betak2 / r2 + lambda2) : betak2 + lambda2;
  int m, n, q;

  pick_ops (M, lambda2, &dvr, &dtr);
  memset (hij, 0, npnp * sizeof (double));
  for (n = 0; n < np; n++) {
    double s = 0.0;
    for (q = 0; q < np; q++) s += Q1[i*np + q] * (dtr[j*np + q] * dtr[n*np + q]);
    hij[RM (i, n, np)] = s;
  }
  for (m = 0; m < np; m++) {
    double s = 0.0;
    for (q = 0; q < np; q++) s += Q2[q*np + j] * (dtr[i*np + q] * dtr[m*np + q]);
    hij[RM (m, j, np)] += s;
  }
  if (M->hasQ3[e])
    for (m = 0; m < np; m++)
      for (n = 0; n < np; n++) {
        hij[RM (m, n, np)] += Q3[RM (i, n, np)] * dvr[RM (n, j, np)] * dvr[RM (i, m, np)];
        hij[RM (m, n, np)] += Q3[RM (m, j, np)] * dvr[RM (j, n, np)] * dvr[RM (m, i, np)];
      }
  hij[RM (i, j, np)] += Q4[RM (i, j, np)] * hCon;
}

void orc_helm_matrix (const orc_mesh* M, int e, double lambda2, double betak2,
                      double* H)
{
  const int np = M->np, npnp = M->npnp;
  int i, j;
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++)
      helm_row (M, e, lambda2, betak2, i, j, H + (i*np + j) * npnp);
}

static int invert (int n, double* A)
/* In-place inverse with partial pivoting (the role of LAPACK dgetrf +
 * dgetri at src/element.cpp:212-216).  Returns nonzero if singular.      */
{
  int i, j, k, *piv = xcalloc (n, sizeof (int));
  double* B = xcalloc (n * n, sizeof (double));
  for (i = 0; i < n; i++) B[i*n + i] = 1.0;
  for (k = 0; k < n; k++) {
    int p = k; double amax = fabs (A[k*n + k]);
    for (i = k + 1; i < n; i++) if (fabs (A[i*n + k]) > amax) { amax = fabs (A[i*n + k]); p = i; }
    if (amax == 0.0) { free (piv); free (B); return 1; }
    if (p != k)
      for (j = 0; j < n; j++) {
        double t = A[k*n + j]; A[k*n + j] = A[p*n + j]; A[p*n + j] = t;
        t = B[k*n + j]; B[k*n + j] = B[p*n + j]; B[p*n + j] = t;
      }
    {
      const double d = 1.0 / A[k*n + k];
      for (j = 0; j < n; j++) { A[k*n + j] *= d; B[k*n + j] *= d; }
    }
    for (i = 0; i < n; i++)
      if (i != k) {
        const double f = A[i*n + k];
        if (f != 0.0)
          for (j = 0; j < n; j++) { A[i*n + j] -= f * A[k*n + j]; B[i*n + j] -= f * B[k*n + j]; }
      }
  }
  memcpy (A, B, n * n * sizeof (double));
  free (piv); free (B);
  return 0;
}

void orc_helm_sc (const orc_mesh* M, int e, double lambda2, double betak2,
                  double* hbb, double* hbi, double* hii)
/* src/element.cpp:131-229 (HelmholtzSC): rows posted in emap order into
 * hbb (next x next), hbi (next x nint), hii (nint x nint); then
 * hii <- hii^-1, hbb <- hbb - hbi hii^-1 hib, hbi <- hbi hii^-1.        */
{
  const int np = M->np, npnp = M->npnp, next = M->next, nint = M->nint;
  double *rmat = xcalloc (npnp, sizeof (double)), *rwrk = xcalloc (npnp, sizeof (double));
  int i, j, k, ij = 0;

  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++, ij++) {
      int eq;
      helm_row (M, e, lambda2, betak2, i, j, rmat);
      for (k = 0; k < npnp; k++) rwrk[k] = rmat[M->emap[k]];
      if ((eq = M->pmap[ij]) < next) {
        memcpy (hbb + eq*next, rwrk, next * sizeof (double));
        memcpy (hbi + eq*nint, rwrk + next, nint * sizeof (double));
      } else
        memcpy (hii + (eq - next)*nint, rwrk + next, nint * sizeof (double));
    }
  if (nint) {
    double* t = xcalloc (next * nint, sizeof (double));
    if (invert (nint, hii)) { fprintf (stderr, "orc: hii singular\n"); abort (); }
    mxm (hbi, next, hii, nint, t, nint);
    for (i = 0; i < next; i++)
      for (j = 0; j < next; j++) {
        double s = 0.0;
        for (k = 0; k < nint; k++) s += hbi[j*nint + k] * t[i*nint + k];
        hbb[i*next + j] -= s;
      }
    memcpy (hbi, t, next * nint * sizeof (double));
    free (t);
  }
  free (rmat); free (rwrk);
}

void orc_helm_diag (const orc_mesh* M, int e, double lambda2, double betak2,
                    double* diag)
/* src/element.cpp:1512-1561 (HelmholtzDiag; cylindrical: HCon =
 * betak2/r^2 + lambda2 per node, 0 where r^2 <= EPSDP).                  */
{
  const int np = M->np, npnp = M->npnp;
  const double *dvr, *dtr;
  const double *Q1 = M->Q1 + e*npnp, *Q2 = M->Q2 + e*npnp, *Q3 = M->Q3 + e*npnp;
  const double *Q4 = M->Q4 + e*npnp;
  const double HCon = lambda2 + betak2;
  double dg[32*32];
  int i, j, q;
  pick_ops (M, lambda2, &dvr, &dtr);
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) {
      const int ij = i*np + j;
      double s = 0.0;
      for (q = 0; q < np; q++) s += Q1[i*np + q] * (dtr[j*np + q] * dtr[j*np + q]);
      dg[ij] = s;
      s = 0.0;
      for (q = 0; q < np; q++) s += Q2[q*np + j] * (dtr[i*np + q] * dtr[i*np + q]);
      dg[ij] += s;
      if (M->hasQ3[e]) dg[ij] += 2.0 * Q3[ij] * dvr[j*np + j] * dvr[i*np + i];
      if (M->cyl) {
        const double r2 = M->y[e*npnp + ij] * M->y[e*npnp + ij];
        dg[ij] += ((r2 > EPSDP) ? (betak2 / r2 + lambda2) : 0.0) * Q4[ij];
      } else
        dg[ij] += HCon * Q4[ij];
    }
  for (i = 0; i < npnp; i++) diag[i] = dg[M->emap[i]];
}

/* ======================================================================== *
 * Matrix systems (src/matrix.cpp:139-323).
 * ======================================================================== */

struct orc_sys {
  const orc_mesh* M;
  double lambda2, betak2;
  int method, singular, nsolve, nband;
  double *H;            /* nband * nsolve, LAPACK 'U' band storage        */
  double *hbi, *hii;    /* nel * next*nint, nel * nint*nint               */
  double *PC;           /* npts                                           */
  int npts;
};

static void band_cholesky (int n, int nband, double* H)
/* Unblocked banded Cholesky, upper storage (the algorithm of LAPACK dpbtf2
 * behind dpbtrf at src/matrix.cpp:259).  U(i,j) at H[j*nband + kd+i-j]. */
{
  const int kd = nband - 1;
  int j, i, l;
  for (j = 0; j < n; j++) {
    double* col = H + j*nband;
    double ajj = col[kd];
    if (ajj <= 0.0) { fprintf (stderr, "orc: band matrix not SPD at %d\n", j); abort (); }
    ajj = sqrt (ajj);
    col[kd] = ajj;
    {
      const int km = (kd < n - 1 - j) ? kd : n - 1 - j;
      /* row j of U: U(j, j+t) at H[(j+t)*nband + kd - t] */
      for (i = 1; i <= km; i++) H[(j+i)*nband + kd - i] /= ajj;
      for (i = 1; i <= km; i++) {
        const double uji = H[(j+i)*nband + kd - i];
        for (l = i; l <= km; l++)
          H[(j+l)*nband + kd - (l - i)] -= uji * H[(j+l)*nband + kd - l];
      }
    }
  }
}

static void band_solve (int n, int nband, const double* H, double* b)
/* dpbtrs 'U' (src/field.cpp:609): U^T y = b, then U x = y.              */
{
  const int kd = nband - 1;
  int j, k;
  for (j = 0; j < n; j++) {
    const double* col = H + j*nband;
    const int k0 = (j - kd > 0) ? j - kd : 0;
    double s = b[j];
    for (k = k0; k < j; k++) s -= col[kd - (j - k)] * b[k];
    b[j] = s / col[kd];
  }
  for (j = n - 1; j >= 0; j--) {
    const int k1 = (j + kd < n - 1) ? j + kd : n - 1;
    double s = b[j];
    for (k = j + 1; k <= k1; k++) s -= H[k*nband + kd - (k - j)] * b[k];
    b[j] = s / H[j*nband + kd];
  }
}

orc_sys* orc_sys_create (const orc_mesh* M, double lambda2, double betak2,
                         int method)
{
  orc_sys* S = xcalloc (1, sizeof (orc_sys));
  const int nel = M->nel, next = M->next, nint = M->nint, npnp = M->npnp;
  int e, i, k;

  S->M = M; S->lambda2 = lambda2; S->betak2 = betak2; S->method = method;
  S->singular = (lambda2 + betak2) < EPSSP && (M->nglobal == M->nsolve);
  S->nsolve = S->singular ? M->nsolve - 1 : M->nsolve;
  S->nband = M->nband;
  S->npts = M->nglobal + nel * nint;

  if (method == ORC_DIRECT) {
    double* hbb = xcalloc (next * next, sizeof (double));
    const int nb = S->nband;
    S->hbi = xcalloc ((size_t) nel * next * (nint ? nint : 1), sizeof (double));
    S->hii = xcalloc ((size_t) nel * (nint ? nint * nint : 1), sizeof (double));
    S->H   = xcalloc ((size_t) nb * (S->nsolve ? S->nsolve : 1), sizeof (double));
    for (e = 0; e < nel; e++) {
      const int* bmap = M->btog + e*next;
      orc_helm_sc (M, e, lambda2, betak2, hbb, S->hbi + (size_t) e*next*nint,
                   S->hii + (size_t) e*nint*nint);
      for (i = 0; i < next; i++) {
        const int m = bmap[i];
        if (m >= S->nsolve) continue;
        for (k = 0; k < next; k++) {
          const int n = bmap[k];
          if (n < S->nsolve && n >= m)
            S->H[(size_t)(n + 1) * nb - n + m - 1] += hbb[i*next + k];  /* band_addr */
        }
      }
    }
    free (hbb);
    if (S->nsolve) band_cholesky (S->nsolve, nb, S->H);
  } else {
    double* ed = xcalloc (npnp, sizeof (double));
    S->PC = xcalloc (S->npts, sizeof (double));
    for (e = 0; e < nel; e++) {
      orc_helm_diag (M, e, lambda2, betak2, ed);
      for (i = 0; i < next; i++) S->PC[M->btog[e*next + i]] += ed[i];
      for (i = 0; i < nint; i++) S->PC[M->nglobal + e*nint + i] = ed[next + i];
    }
    for (i = 0; i < S->npts; i++) S->PC[i] = 1.0 / S->PC[i];
    free (ed);
  }
  return S;
}

void orc_sys_destroy (orc_sys* S)
{
  if (!S) return;
  free (S->H); free (S->hbi); free (S->hii); free (S->PC); free (S);
}

void orc_sys_info (const orc_sys* S, int* info)
{ info[0] = S->nsolve; info[1] = S->nband; info[2] = S->singular; }

void orc_sys_band (const orc_sys* S, double* H)
{ if (S->H) memcpy (H, S->H, (size_t) S->nband * S->nsolve * sizeof (double)); }

/* --- Field::solve pieces ------------------------------------------------- */

static void get_essential (const orc_mesh* M, const double* bc, double* tgt)
/* src/field.cpp:994-1019 with Essential::set (src/condition.cpp:80-101). */
{
  const int np = M->np, nm = np - 1, next = M->next;
  int b, i;
  for (b = 0; b < M->nbound; b++) {
    const int* bmap = M->btog + M->belmt[b] * next;
    const int  side = M->bside[b];
    const int* start = bmap + side * nm;
    const double* src = bc + b * np;
    if (M->bkind[b] != 1) continue;
    for (i = 0; i < nm; i++) tgt[start[i]] = src[i];
    if (side == 3) tgt[bmap[0]] = src[nm];
    else           tgt[start[nm]] = src[nm];
  }
}

static void natural_sum (const orc_mesh* M, const double* bc, double* tgt)
/* <h, w> (src/field.cpp:911-918) with Natural::sum (condition.cpp:192-219). */
{
  const int np = M->np, nm = np - 1, next = M->next;
  int b, i;
  for (b = 0; b < M->nbound; b++) {
    const int* bmap = M->btog + M->belmt[b] * next;
    const int  side = M->bside[b];
    const int* start = bmap + side * nm;
    const double* src = bc + b * np;
    const double* wt = M->barea + b * np;
    if (M->bkind[b] != 2) continue;
    for (i = 0; i < nm; i++) tgt[start[i]] += src[i] * wt[i];
    if (side == 3) tgt[bmap[0]] += src[nm] * wt[nm];
    else           tgt[start[nm]] += src[nm] * wt[nm];
  }
}

static void constrain (const orc_sys* S, double* force, const double* esstlbc)
/* src/field.cpp:728-768: force <- -(M f + H g).                         */
{
  const orc_mesh* M = S->M;
  const int nel = M->nel, npnp = M->npnp, next = M->next;
  double u[32*32];
  int e, i;
  for (e = 0; e < nel; e++) {
    double* f = force + e*npnp;
    const double* Q8 = M->Q8 + e*npnp;    /* Element::weight: not times y */
    for (i = 0; i < npnp; i++) f[i] *= Q8[i];
    if (M->emask[e]) {
      const int* btog = M->btog + e*next;
      memset (u, 0, npnp * sizeof (double));
      for (i = 0; i < next; i++) u[M->emap[i]] = esstlbc[btog[i]];
      orc_helm_op (M, e, S->lambda2, S->betak2, u, u);
      for (i = 0; i < npnp; i++) f[i] += u[i];
    }
  }
  for (i = 0; i < M->nPlane; i++) force[i] = -force[i];
}

static void set_essential (const orc_mesh* M, const double* src, double* tgt)
/* src/field.cpp:1022-1040 with Element::bndryMask (element.cpp:500-533). */
{
  const int nel = M->nel, npnp = M->npnp, next = M->next;
  int e, i;
  for (e = 0; e < nel; e++) {
    if (!M->emask[e]) continue;
    for (i = 0; i < next; i++)
      if (M->bmask[e*next + i])
        tgt[e*npnp + M->emap[i]] = src[M->btog[e*next + i]];
  }
}

void orc_helm_global (const orc_sys* S, const double* x, double* y)
/* src/field.cpp:771-856 (Field::HelmholtzOperator), scalar branch.      */
{
  const orc_mesh* M = S->M;
  const int nel = M->nel, next = M->next, nint = M->nint;
  double P[32*32];
  int e, i;
  memset (y, 0, S->npts * sizeof (double));
  for (e = 0; e < nel; e++) {
    const int* gid = M->btog + e*next;
    const double* xint = x + M->nglobal + e*nint;
    double* yint = y + M->nglobal + e*nint;
    for (i = 0; i < next; i++) P[M->emap[i]] = x[gid[i]];
    for (i = 0; i < nint; i++) P[M->emap[next + i]] = xint[i];
    orc_helm_op (M, e, S->lambda2, S->betak2, P, P);
    for (i = 0; i < next; i++) y[gid[i]] += P[M->emap[i]];
    for (i = 0; i < nint; i++) yint[i] += P[M->emap[next + i]];
  }
}

static double dot (int n, const double* a, const double* b)
{
  double s = 0.0; int i;
  for (i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

int orc_solve_plane (const orc_sys* S, double* forcing, const double* bc,
                     double* unknown, double tol_rel, int step_max)
/* Field::solve for one plane, src/field.cpp:509-725.                     */
{
  const orc_mesh* M = S->M;
  const int nel = M->nel, npnp = M->npnp, next = M->next, nint = M->nint;
  const int nglobal = M->nglobal, nsolve = S->nsolve, nzero = nglobal - nsolve;
  int e, i;

  if (S->method == ORC_DIRECT) {
    double* RHS = xcalloc (nglobal, sizeof (double));
    double* work = xcalloc (npnp, sizeof (double));
    get_essential (M, bc, RHS);
    constrain (S, forcing, RHS);
    /* buildRHS (field.cpp:859-925) with local2globalSumSC (element.cpp:356-395) */
    memset (RHS, 0, nglobal * sizeof (double));
    for (e = 0; e < nel; e++) {
      double* F = forcing + e*npnp;
      const double* hbi = S->hbi + (size_t) e*next*nint;
      const int* btog = M->btog + e*next;
      int b, k;
      for (i = 0; i < npnp; i++) work[i] = F[M->emap[i]];
      memcpy (F, work, npnp * sizeof (double));
      for (b = 0; b < next; b++) {
        double s = 0.0;
        for (k = 0; k < nint; k++) s += hbi[b*nint + k] * F[next + k];
        F[b] -= s;
      }
      for (b = 0; b < next; b++) RHS[btog[b]] += F[b];
    }
    natural_sum (M, bc, RHS);
    memset (RHS + nsolve, 0, nzero * sizeof (double));
    if (nsolve) band_solve (nsolve, S->nband, S->H, RHS);
    /* global2localSC (element.cpp:398-438) */
    for (e = 0; e < nel; e++) {
      double* F = forcing + e*npnp;
      double* tgt = unknown + e*npnp;
      const double* hbi = S->hbi + (size_t) e*next*nint;
      const double* hii = S->hii + (size_t) e*nint*nint;
      const int* btog = M->btog + e*next;
      double wint[32*32], ub[128];
      int b, k;
      for (b = 0; b < next; b++) { ub[b] = RHS[btog[b]]; tgt[M->emap[b]] = ub[b]; }
      for (k = 0; k < nint; k++) wint[k] = F[next + k];
      for (k = 0; k < nint; k++) {
        double s = 0.0; int c;
        for (c = 0; c < nint; c++) s += hii[k*nint + c] * wint[c];
        for (b = 0; b < next; b++) s -= hbi[b*nint + k] * ub[b];
        tgt[M->emap[next + k]] = s;
      }
    }
    memset (RHS, 0, nglobal * sizeof (double));
    get_essential (M, bc, RHS);
    set_essential (M, RHS, unknown);
    free (RHS); free (work);
    return 0;
  } else {
    const int npts = S->npts;
    double *r = xcalloc (npts, sizeof (double)), *p = xcalloc (npts, sizeof (double));
    double *q = xcalloc (npts, sizeof (double)), *x = xcalloc (npts, sizeof (double));
    double *z = xcalloc (npts, sizeof (double));
    double epsb2, r2, rho1, rho2 = 1.0, alpha, beta;
    int it = 0;

    get_essential (M, bc, x);
    constrain (S, forcing, x);
    /* buildRHS iterative: local2globalSum */
    for (e = 0; e < nel; e++) {
      const double* F = forcing + e*npnp;
      const int* btog = M->btog + e*next;
      double* rint = r + nglobal + e*nint;
      for (i = 0; i < next; i++) r[btog[i]] += F[M->emap[i]];
      for (i = 0; i < nint; i++) rint[i] += F[M->emap[next + i]];
    }
    natural_sum (M, bc, r);
    memset (r + nsolve, 0, nzero * sizeof (double));

    epsb2 = tol_rel * sqrt (dot (npts, r, r));
    epsb2 *= epsb2;

    /* local2global (field.cpp:928-947): x from the element store          */
    for (e = 0; e < nel; e++) {
      const double* src = unknown + e*npnp;
      const int* btog = M->btog + e*next;
      for (i = 0; i < next; i++) x[btog[i]] = src[M->emap[i]];
      for (i = 0; i < nint; i++) x[nglobal + e*nint + i] = src[M->emap[next + i]];
    }
    memset (x + nsolve, 0, nzero * sizeof (double));
    memcpy (q, x, npts * sizeof (double));
    orc_helm_global (S, q, p);
    memset (p + nsolve, 0, nzero * sizeof (double));
    memset (r + nsolve, 0, nzero * sizeof (double));
    for (i = 0; i < npts; i++) r[i] -= p[i];
    r2 = dot (npts, r, r);

    while (r2 > epsb2 && ++it < step_max) {
      for (i = 0; i < npts; i++) z[i] = S->PC[i] * r[i];
      rho1 = dot (npts, r, z);
      if (it == 1) memcpy (p, z, npts * sizeof (double));
      else {
        beta = rho1 / rho2;
        for (i = 0; i < npts; i++) p[i] = z[i] + beta * p[i];
      }
      orc_helm_global (S, p, q);
      memset (q + nsolve, 0, nzero * sizeof (double));
      alpha = rho1 / dot (npts, p, q);
      for (i = 0; i < npts; i++) x[i] += alpha * p[i];
      for (i = 0; i < npts; i++) r[i] -= alpha * q[i];
      rho2 = rho1;
      r2 = dot (npts, r, r);
    }
    /* global2local (field.cpp:950-969) */
    for (e = 0; e < nel; e++) {
      double* tgt = unknown + e*npnp;
      const int* btog = M->btog + e*next;
      for (i = 0; i < next; i++) tgt[M->emap[i]] = x[btog[i]];
      for (i = 0; i < nint; i++) tgt[M->emap[next + i]] = x[nglobal + e*nint + i];
    }
    get_essential (M, bc, x);
    set_essential (M, x, unknown);
    free (r); free (p); free (q); free (x); free (z);
    return (it == step_max) ? -1 : it;
  }
}

static int field_solve_planes (const orc_mesh* M, int nz, double lambda2, double beta,
                        int base, int method, double* f, const double* bc, double* u,
                        double tol_rel, int step_max, int nthreads, int nrep,
                        double* t_solve, int* iters)
/* Field::solve over planes (field.cpp:564-723) with one MatrixSys per
 * mode base..base+nz/2-1 (matrix.cpp:33-110; betak2 with the z SVV
 * weights of orc_set_svv_z, 1 unless set).  The Nyquist plane 1 is skipped only on the planes that start at
 * mode 0 (the ROOTONLY skip of field.cpp:566).  iters (may be NULL)
 * receives each plane's PCG iteration count, -1 at the limit.            */
{
  const int nmodes = (nz > 1) ? nz / 2 : 1;
  const int nline = M->nbound * M->np;
  orc_sys** sys = xcalloc (nmodes, sizeof (orc_sys*));
  int m, k, warn = 0;
  double t0;

  double* Sz = xcalloc (nmodes, sizeof (double));
  orc_svv_coeffs_z (nmodes, svvz_nmode ? svvz_nmode : base + nmodes, base, svvz_mz, svvz_eps,
                    svvz_nu, Sz);
  if (nthreads > 0) omp_set_num_threads (nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (m = 0; m < nmodes; m++) {
    /* MIXED: DIRECT for mode 0, JACPCG otherwise (matrix.cpp:93-96) */
    const int kind = (method == 2) ? (base + m == 0 ? ORC_DIRECT : ORC_JACPCG) : method;
    const double bk = mode_constant (M, base + m, beta);
    /* betak2 weighted by the z SVV coefficient for lambda2 > EPSDP
     * (src/matrix.cpp:78-83)                                            */
    const double bk2 = (lambda2 > EPSDP) ? bk * bk * Sz[m] : bk * bk;
    sys[m] = orc_sys_create (M, lambda2, bk2, kind);
  }
  free (Sz);
  {
    /* repeated solves restore the (destroyed) forcing from a copy, outside
     * the timed loop body's plane solves                                  */
    const size_t ntot = (size_t) nz * M->planeSize;
    double* f0 = (nrep > 1) ? xcalloc (ntot, sizeof (double)) : NULL;
    double tsum = 0.0;
    int rep;
    if (f0) memcpy (f0, f, ntot * sizeof (double));
    if (nrep < 1) nrep = 1;
    for (rep = 0; rep < nrep; rep++) {
      if (rep) memcpy (f, f0, ntot * sizeof (double));
      t0 = omp_get_wtime ();
#pragma omp parallel for schedule(dynamic, 1) reduction(+:warn)
      for (k = 0; k < nz; k++) {
        int it;
        if (nz > 1 && k == 1 && base == 0) continue;
        it = orc_solve_plane (sys[k >> 1], f + (size_t) k * M->planeSize,
                              bc + (size_t) k * nline, u + (size_t) k * M->planeSize,
                              tol_rel, step_max);
        if (it < 0) warn++;
        if (iters) iters[k] = it;
      }
      tsum += omp_get_wtime () - t0;
    }
    if (t_solve) *t_solve = tsum / nrep;
    free (f0);
  }
  for (m = 0; m < nmodes; m++) orc_sys_destroy (sys[m]);
  free (sys);
  return warn;
}

int orc_field_solve (const orc_mesh* M, int nz, double lambda2, double beta,
                     int method, double* f, const double* bc, double* u,
                     double tol_rel, int step_max, int nthreads, int nrep,
                     double* t_solve)
{
  return field_solve_planes (M, nz, lambda2, beta, 0, method, f, bc, u, tol_rel, step_max,
                      nthreads, nrep, t_solve, NULL);
}

int orc_field_solve_iters (const orc_mesh* M, int nz, double lambda2, double beta,
                           int method, double* f, const double* bc, double* u,
                           double tol_rel, int step_max, int* iters)
{
  return field_solve_planes (M, nz, lambda2, beta, 0, method, f, bc, u, tol_rel, step_max,
                      0, 1, NULL, iters);
}

int orc_field_solve_base (const orc_mesh* M, int nz, double lambda2, double beta,
                          int base, int method, double* f, const double* bc, double* u,
                          double tol_rel, int step_max, int* iters)
{
  return field_solve_planes (M, nz, lambda2, beta, base, method, f, bc, u, tol_rel,
                      step_max, 0, 1, NULL, iters);
}

/* ======================================================================== *
 * Fourier transform: dDFTr semantics (femlib/fourier.c:47-171, packing of
 * femlib/temfftd.F:161-274).  The complex FFT itself is a plain recursive
 * mixed-radix (2,3,5) Cooley-Tukey, not Temperton's GPFA.
 * ======================================================================== */

typedef struct { double re, im; } cplx;

static void fft_rec (cplx* x, cplx* scratch, int n, int stride, int sign)
/* In-place DFT of x[0], x[stride], ... (n points), exponent sign*2pi i/n. */
{
  int p, m, q, s, r;
  if (n == 1) return;
  p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : 5;
  m = n / p;
  /* decimate: sub-sequence r = x[(q*p + r)*stride] */
  for (r = 0; r < p; r++)
    for (q = 0; q < m; q++) scratch[r*m + q] = x[(q*p + r) * stride];
  for (r = 0; r < p; r++) fft_rec (scratch + r*m, x, m, 1, sign);
  for (q = 0; q < m; q++)
    for (s = 0; s < p; s++) {
      const int kk = q + s*m;
      double re = 0.0, im = 0.0;
      for (r = 0; r < p; r++) {
        const double ang = sign * 2.0 * M_PI * (double) ((long) r * kk % n) / n;
        const double c = cos (ang), sn = sin (ang);
        const cplx v = scratch[r*m + q];
        re += v.re * c - v.im * sn;
        im += v.re * sn + v.im * c;
      }
      x[kk * stride].re = re; x[kk * stride].im = im;
    }
}

int orc_dftr (double* data, int tlen, int ntrn, int sign)
{
  const int nph = ntrn / 2, nzh = tlen / 2;
  int t = tlen, i, j;
  if (tlen < 2 || !ntrn) return 0;
  while (t % 2 == 0) t /= 2;
  while (t % 3 == 0) t /= 3;
  while (t % 5 == 0) t /= 5;
  if (t != 1 || (ntrn & 1)) return -1;

#pragma omp parallel for private(j)
  for (i = 0; i < nph; i++) {
    cplx* c = xcalloc (tlen, sizeof (cplx));
    cplx* s = xcalloc (tlen, sizeof (cplx));
    const int ii = i + nph;
    if (sign == 1) {
      for (j = 0; j < tlen; j++) {
        c[j].re = data[(size_t) j*ntrn + i];
        c[j].im = data[(size_t) j*ntrn + ii];
      }
      fft_rec (c, s, tlen, 1, -1);
      for (j = 1; j < nzh; j++) {
        const int mj = tlen - j;
        data[(size_t)(2*j)*ntrn + i]    = 0.5 * (c[j].re + c[mj].re) / tlen;
        data[(size_t)(2*j+1)*ntrn + i]  = 0.5 * (c[j].im - c[mj].im) / tlen;
        data[(size_t)(2*j)*ntrn + ii]   = 0.5 * (c[j].im + c[mj].im) / tlen;
        data[(size_t)(2*j+1)*ntrn + ii] = 0.5 * (c[mj].re - c[j].re) / tlen;
      }
      data[i]               = c[0].re / tlen;
      data[(size_t) ntrn + i]  = c[nzh].re / tlen;
      data[ii]              = c[0].im / tlen;
      data[(size_t) ntrn + ii] = c[nzh].im / tlen;
    } else {
      for (j = 1; j < nzh; j++) {
        const int mj = tlen - j;
        const double are = data[(size_t)(2*j)*ntrn + i], aim = data[(size_t)(2*j+1)*ntrn + i];
        const double bre = data[(size_t)(2*j)*ntrn + ii], bim = data[(size_t)(2*j+1)*ntrn + ii];
        c[j].re  = are - bim;  c[j].im  = aim + bre;
        c[mj].re = are + bim;  c[mj].im = bre - aim;
      }
      c[0].re = data[i];   c[0].im = data[ii];
      c[nzh].re = data[(size_t) ntrn + i]; c[nzh].im = data[(size_t) ntrn + ii];
      fft_rec (c, s, tlen, 1, +1);
      for (j = 0; j < tlen; j++) {
        data[(size_t) j*ntrn + i]  = c[j].re;
        data[(size_t) j*ntrn + ii] = c[j].im;
      }
    }
    free (c); free (s);
  }
  return 0;
}

/* ======================================================================== *
 * Field operators.
 * ======================================================================== */

static void elmt_grad (const orc_mesh* M, int e, double* tgt, int dir)
/* src/element.cpp:552-597 (Element::grad), plain DV (no SVV).           */
{
  const int np = M->np, npnp = M->npnp;
  const double* rx = dir ? M->drdy + e*npnp : M->drdx + e*npnp;
  const double* sx = dir ? M->dsdy + e*npnp : M->dsdx + e*npnp;
  const int hr = dir ? M->hasdrdy[e] : M->hasdrdx[e];
  const int hs = dir ? M->hasdsdy[e] : M->hasdsdx[e];
  double A[32*32], B[32*32];
  int i;
  if (hr && hs) {
    mxm (tgt, np, M->DT, np, A, np);
    mxm (M->DV, np, tgt, np, B, np);
    for (i = 0; i < npnp; i++) tgt[i] = B[i] * sx[i] + A[i] * rx[i];
  } else if (hr) {
    mxm (tgt, np, M->DT, np, A, np);
    for (i = 0; i < npnp; i++) tgt[i] = A[i] * rx[i];
  } else {
    mxm (M->DV, np, tgt, np, B, np);
    for (i = 0; i < npnp; i++) tgt[i] = B[i] * sx[i];
  }
}

void orc_gradient (const orc_mesh* M, int nz, double beta, double* data,
                   int dir)
/* src/auxfield.cpp:514-665 (AuxField::gradient), scalar branch.         */
{
  const int nP = M->planeSize;
  int k, e;
  if (dir < 2) {
#pragma omp parallel for private(e)
    for (k = 0; k < nz; k++)
      for (e = 0; e < M->nel; e++)
        elmt_grad (M, e, data + (size_t) k*nP + e*M->npnp, dir);
  } else {
    const int nmodes = nz / 2;
    memset (data, 0, 2 * (size_t) nP * sizeof (double));
#pragma omp parallel for
    for (k = 1; k < nmodes; k++) {
      double* re = data + (size_t)(2*k) * nP;
      double* im = re + nP;
      const double bk = beta * k;
      int i;
      for (i = 0; i < nP; i++) {
        const double t = re[i];
        re[i] = -bk * im[i];
        im[i] =  bk * t;
      }
    }
  }
}

void orc_smooth (const orc_mesh* M, int nz, double* data)
/* src/field.cpp:190-236 (Field::smooth) with bndryDsSum / bndryInsert
 * (src/element.cpp:464-497, 536-549).                                    */
{
  const int nP = M->planeSize, next = M->next, npnp = M->npnp;
  int k;
#pragma omp parallel for
  for (k = 0; k < nz; k++) {
    double* src = data + (size_t) k * nP;
    double* dssum = xcalloc (M->nglobal, sizeof (double));
    int e, i;
    for (e = 0; e < M->nel; e++)
      for (i = 0; i < next; i++) {
        const int q = M->emap[i];
        dssum[M->btog[e*next + i]] += M->Q4[e*npnp + q] * src[e*npnp + q];
      }
    for (i = 0; i < M->nglobal; i++) dssum[i] *= M->imass[i];
    for (e = 0; e < M->nel; e++)
      for (i = 0; i < next; i++)
        src[e*npnp + M->emap[i]] = dssum[M->btog[e*next + i]];
    free (dssum);
  }
}

/* ======================================================================== *
 * DNS step (dns/integrate.cpp:79-395, dns/nonlinear.cpp:36-411),
 * Cartesian coordinates, no boundary conditions, no forcing, no scalar.
 * ======================================================================== */

struct orc_dns {
  const orc_mesh* M;        /* velocity mesh (its essential sides: walls)  */
  const orc_mesh* Mp;       /* pressure mesh (its natural sides: computed
                             * high-order BCs); M when there are none      */
  int nz, ncom, nord, advection, step, nthreads, step_max, toggle;
  double beta, dt, kinvis, tol_rel, time;
  size_t ntot;
  double **Us, **Uf;        /* [nord][ncom] level pointers                 */
  double *store;
  orc_sys **vsys, **psys;   /* per mode                                   */
  /* boundaries (M->nbound > 0): velocity BC lines (Fourier space), the
   * pressure BC lines and the computed-BC history of BCmgr
   * (src/bcmgr.cpp:920-996): u.n and the high-order pressure flux per
   * time level                                                           */
  int nline;
  double *bcu[3], *bcp;
  double **un, **hop;       /* [nord] level pointers, nz*nline each        */
  double *bstore;
};

static void stiffly_stable (int n, double* c)
/* src/integration.cpp:99-134 */
{
  switch (n) {
  case 1: c[0] = 1.0; c[1] = -1.0; break;
  case 2: c[0] = 1.5; c[1] = -2.0; c[2] = 0.5; break;
  default: c[0] = 11.0/6.0; c[1] = -3.0; c[2] = 1.5; c[3] = -1.0/3.0; break;
  }
}

static void extrapolation (int n, double* c)
/* src/integration.cpp:137-164 */
{
  switch (n) {
  case 1: c[0] = 1.0; break;
  case 2: c[0] = 2.0; c[1] = -1.0; break;
  default: c[0] = 3.0; c[1] = -3.0; c[2] = 1.0; break;
  }
}

orc_dns* orc_dns_create (const orc_mesh* M, int nz, double beta, double dt,
                         double kinvis, int norder, int advection,
                         double tol_rel, int step_max, int nthreads)
{
  orc_dns* D = xcalloc (1, sizeof (orc_dns));
  const int nmodes = (nz > 1) ? nz / 2 : 1;
  double alpha[4];
  int q, i, m;
  D->M = M; D->Mp = M; D->nz = nz; D->nord = norder; D->advection = advection;
  D->beta = beta; D->dt = dt; D->kinvis = kinvis; D->tol_rel = tol_rel;
  D->step_max = step_max; D->nthreads = nthreads; D->toggle = 1;
  D->ncom = (nz > 1) ? 3 : 2;
  D->ntot = (size_t) nz * M->planeSize;
  D->store = xcalloc (2 * (size_t) norder * D->ncom * D->ntot, sizeof (double));
  D->Us = xcalloc (norder * D->ncom, sizeof (double*));
  D->Uf = xcalloc (norder * D->ncom, sizeof (double*));
  for (q = 0; q < norder; q++)
    for (i = 0; i < D->ncom; i++) {
      D->Us[q*D->ncom + i] = D->store + (2 * (size_t)(q*D->ncom + i)) * D->ntot;
      D->Uf[q*D->ncom + i] = D->store + (2 * (size_t)(q*D->ncom + i) + 1) * D->ntot;
    }
  stiffly_stable (norder, alpha);
  D->vsys = xcalloc (nmodes, sizeof (orc_sys*));
  D->psys = xcalloc (nmodes, sizeof (orc_sys*));
  if (nthreads > 0) omp_set_num_threads (nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (m = 0; m < nmodes; m++) {
    const double bk = beta * m;
    D->vsys[m] = orc_sys_create (M, alpha[0] / (dt * kinvis), bk * bk, ORC_DIRECT);
    D->psys[m] = orc_sys_create (M, 0.0, bk * bk, ORC_DIRECT);
  }
  return D;
}

/* -- Computed pressure boundary conditions (KIO91) ---------------------- */

static int side_node (int np, int s, int q)
/* element-storage index of node q (CCW) of side s (Edge offsets and
 * skips, src/edge.cpp:43-50)                                            */
{
  switch (s) {
  case 0:  return q;
  case 1:  return q*np + np - 1;
  case 2:  return (np - 1)*np + (np - 1 - q);
  default: return (np - 1 - q)*np;
  }
}

static void side_normal (const orc_mesh* M, int e, int s, double* nx, double* ny)
/* unit outward normal at the side's nodes, CCW (Element::sideGeom,
 * src/element.cpp:1653-1744)                                            */
{
  const int np = M->np;
  int q;
  for (q = 0; q < np; q++) {
    const int i = e*M->npnp + side_node (np, s, q);
    double a, b, len;
    switch (s) {
    case 0:  a = -M->dsdx[i]; b = -M->dsdy[i]; break;
    case 1:  a =  M->drdx[i]; b =  M->drdy[i]; break;
    case 2:  a =  M->dsdx[i]; b =  M->dsdy[i]; break;
    default: a = -M->drdx[i]; b = -M->drdy[i]; break;
    }
    len = hypot (a, b);
    nx[q] = a / len; ny[q] = b / len;
  }
}

static void side_grad (const orc_mesh* M, int e, int s, const double* src, double* gx,
                       double* gy)
/* x and y derivatives of element data src at the side's nodes, CCW
 * (Element::sideGrad, src/element.cpp:652-714)                           */
{
  const int np = M->np;
  int q, k;
  for (q = 0; q < np; q++) {
    const int idx = side_node (np, s, q), i = idx / np, j = idx % np;
    const int g = e*M->npnp + idx;
    double ddr = 0.0, dds = 0.0;
    for (k = 0; k < np; k++) {
      ddr += M->DV[j*np + k] * src[i*np + k];
      dds += M->DV[i*np + k] * src[k*np + j];
    }
    if (gx) gx[q] = ddr * M->drdx[g] + dds * M->dsdx[g];
    if (gy) gy[q] = ddr * M->drdy[g] + dds * M->dsdy[g];
  }
}

static void curl_curl (const orc_mesh* M, int e, int s, double betaK, const double* Ur,
                       const double* Ui, const double* Vr, const double* Vi,
                       const double* Wr, const double* Wi, double* xr, double* xi,
                       double* yr, double* yi)
/* Fourier-mode curl curl u along a side, first two components
 * (Edge::curlCurl, src/edge.cpp:98-258, Cartesian): for betaK == 0 the
 * 2D / mode-0 form from the real planes only.                          */
{
  const int np = M->np, npnp = M->npnp;
  double w[32*32], a[32*32], b[32*32], t[32];
  int i, q;
  /* real parts: z vorticity w = dV/dx - dU/dy, then (dw/dy, -dw/dx) */
  memcpy (a, Vr, npnp * sizeof (double)); elmt_grad (M, e, a, 0);
  memcpy (b, Ur, npnp * sizeof (double)); elmt_grad (M, e, b, 1);
  for (i = 0; i < npnp; i++) w[i] = a[i] - b[i];
  side_grad (M, e, s, w, yr, xr);
  for (q = 0; q < np; q++) yr[q] = -yr[q];
  if (betaK == 0.0) return;
  memcpy (a, Vi, npnp * sizeof (double)); elmt_grad (M, e, a, 0);
  memcpy (b, Ui, npnp * sizeof (double)); elmt_grad (M, e, b, 1);
  for (i = 0; i < npnp; i++) w[i] = a[i] - b[i];
  side_grad (M, e, s, w, yi, xi);
  for (q = 0; q < np; q++) yi[q] = -yi[q];
  /* semi-Fourier terms of W */
  memcpy (a, Wr, npnp * sizeof (double)); elmt_grad (M, e, a, 0);
  memcpy (b, Wr, npnp * sizeof (double)); elmt_grad (M, e, b, 1);
  for (q = 0; q < np; q++) {
    const int n = side_node (np, s, q);
    xi[q] += betaK * a[n];
    yi[q] += betaK * b[n];
  }
  memcpy (a, Wi, npnp * sizeof (double)); elmt_grad (M, e, a, 0);
  memcpy (b, Wi, npnp * sizeof (double)); elmt_grad (M, e, b, 1);
  for (q = 0; q < np; q++) {
    const int n = side_node (np, s, q);
    xr[q] -= betaK * a[n];
    yr[q] -= betaK * b[n];
  }
  /* second derivatives in z */
  for (q = 0; q < np; q++) {
    const int n = side_node (np, s, q);
    xr[q] += betaK * betaK * Ur[n];
    xi[q] += betaK * betaK * Ui[n];
    yr[q] += betaK * betaK * Vr[n];
    yi[q] += betaK * betaK * Vi[n];
  }
  (void) t;
}

static void rollv (double** lev, int n)
/* rollv (src/sem.h:90-103)                                              */
{
  double* t;
  int q;
  if (n < 2) return;
  t = lev[n - 1];
  for (q = n - 1; q; q--) lev[q] = lev[q - 1];
  lev[0] = t;
}

static void maintain_fourier (orc_dns* D, double** Us, double** Uf)
/* BCmgr::maintainFourier (src/bcmgr.cpp:1039-1298), the terms of the
 * high-order pressure BC: at each boundary node of every plane,
 *   hop = -(1/dt)(a0 u.n + sum_q a_{q+1} (u.n)_q)   (Je = min(step-1,
 *         N_TIME) > 0, the velocity on the boundary may vary in time)
 *       + n.N(u) - nu n.curlCurl(u),
 * with u.n of past levels kept; velocity in Fourier space (Us[0]),
 * nonlinear terms N (Uf[0]).  Plane 1 (Nyquist) is skipped.  The
 * div(u) and grad(u).n stores of the reference serve other BC kinds and
 * are not formed.                                                       */
{
  const orc_mesh* M = D->Mp;
  const int np = M->np, npnp = M->npnp, nz = D->nz, nline = D->nline, ps = M->planeSize;
  const int Je = (D->step - 1 < D->nord) ? D->step - 1 : D->nord;
  const double nu = D->kinvis, invDt = 1.0 / D->dt;
  double alpha[4], nx[32], ny[32], xr[32], xi[32], yr[32], yi[32];
  double* unew = xcalloc ((size_t) nz * nline, sizeof (double));
  int b, k, q, m;
  if (Je) stiffly_stable (Je, alpha);
  rollv (D->hop, D->nord);
  memset (D->hop[0], 0, (size_t) nz * nline * sizeof (double));
  for (b = 0; b < M->nbound; b++) {
    const int e = M->belmt[b], s = M->bside[b], j = b * np;
    side_normal (M, e, s, nx, ny);
    for (k = 0; k < nz; k++) {
      const double *ux, *uy, *fx, *fy;
      double* h = D->hop[0] + (size_t) k * nline + j;
      double* un = unew + (size_t) k * nline + j;
      if (nz > 1 && k == 1) continue;
      ux = Us[0] + (size_t) k * ps + e * npnp;
      uy = Us[1] + (size_t) k * ps + e * npnp;
      fx = Uf[0] + (size_t) k * ps + e * npnp;
      fy = Uf[1] + (size_t) k * ps + e * npnp;
      for (q = 0; q < np; q++) {
        const int n = side_node (np, s, q);
        un[q] = ux[n] * nx[q] + uy[n] * ny[q];
      }
      if (Je) {
        int l;
        for (q = 0; q < np; q++) {
          double wk = alpha[0] * un[q];
          for (l = 0; l < Je; l++) wk += alpha[l + 1] * D->un[l][(size_t) k * nline + j + q];
          h[q] = -invDt * wk;
        }
      }
      for (q = 0; q < np; q++) {
        const int n = side_node (np, s, q);
        h[q] += fx[n] * nx[q] + fy[n] * ny[q];
      }
    }
  }
  rollv (D->un, D->nord);
  memcpy (D->un[0], unew, (size_t) nz * nline * sizeof (double));
  free (unew);
  for (b = 0; b < M->nbound; b++) {
    const int e = M->belmt[b], s = M->bside[b], j = b * np;
    side_normal (M, e, s, nx, ny);
    {
      const double* ur = Us[0] + e * npnp;
      const double* vr = Us[1] + e * npnp;
      double* h = D->hop[0] + j;
      curl_curl (M, e, s, 0.0, ur, NULL, vr, NULL, NULL, NULL, xr, NULL, yr, NULL);
      for (q = 0; q < np; q++) h[q] += -nu * xr[q] * nx[q] + -nu * yr[q] * ny[q];
    }
    for (m = 1; m < (nz > 1 ? nz / 2 : 1); m++) {
      const size_t kr = 2 * (size_t) m, ki = kr + 1;
      const double* Ur = Us[0] + kr * ps + e * npnp; const double* Ui = Us[0] + ki * ps + e * npnp;
      const double* Vr = Us[1] + kr * ps + e * npnp; const double* Vi = Us[1] + ki * ps + e * npnp;
      const double* Wr = Us[2] + kr * ps + e * npnp; const double* Wi = Us[2] + ki * ps + e * npnp;
      double* hr = D->hop[0] + kr * nline + j;
      double* hi = D->hop[0] + ki * nline + j;
      curl_curl (M, e, s, m * D->beta, Ur, Ui, Vr, Vi, Wr, Wi, xr, xi, yr, yi);
      for (q = 0; q < np; q++) {
        hr[q] += -nu * xr[q] * nx[q];
        hi[q] += -nu * xi[q] * nx[q];
        hr[q] += -nu * yr[q] * ny[q];
        hi[q] += -nu * yi[q] * ny[q];
      }
    }
  }
}

static void evaluate_pressure_bcs (orc_dns* D)
/* Field::evaluateBoundaries of the pressure with BCmgr::evaluateCNBCp
 * (src/bcmgr.cpp:1301-1337) on every natural side of the pressure mesh:
 * the flux at the end of the step extrapolated from the stored levels,
 * sum_{q < Je} beta_q hop_q, Je = min(step, N_TIME); essential sides of
 * the pressure (p = 0) and plane 1 zero.                                */
{
  const orc_mesh* M = D->Mp;
  const int np = M->np, nz = D->nz, nline = D->nline;
  const int Je = (D->step < D->nord) ? D->step : D->nord;
  double beta[3];
  int b, k, q, l;
  extrapolation (Je, beta);
  memset (D->bcp, 0, (size_t) nz * nline * sizeof (double));
  for (b = 0; b < M->nbound; b++) {
    if (M->bkind[b] != 2) continue;
    for (k = 0; k < nz; k++) {
      if (nz > 1 && k == 1) continue;
      for (q = 0; q < np; q++) {
        const size_t i = (size_t) k * nline + b * np + q;
        double v = 0.0;
        for (l = 0; l < Je; l++) v += beta[l] * D->hop[l][i];
        D->bcp[i] = v;
      }
    }
  }
}

orc_dns* orc_dns_create_bc (const orc_mesh* M, const orc_mesh* Mp, int nz, double beta,
                            double dt, double kinvis, int norder, int advection,
                            double tol_rel, int step_max, int nthreads, const double* bcu,
                            const double* bcv, const double* bcw)
/* A DNS on a mesh with boundaries: M numbers the velocity (its essential
 * sides carry the given Fourier-space BC lines bcu/bcv/bcw, nz*nline
 * each, static in time), Mp the pressure (same elements and boundary
 * list; its natural sides take the computed high-order flux, its
 * essential sides p = 0).                                               */
{
  orc_dns* D = orc_dns_create (M, nz, beta, dt, kinvis, norder, advection, tol_rel,
                               step_max, nthreads);
  const int nmodes = (nz > 1) ? nz / 2 : 1;
  const size_t nl = (size_t) nz * M->nbound * M->np;
  const double* src[3] = {bcu, bcv, bcw};
  int m, i, q;
  D->Mp = Mp;
  D->nline = M->nbound * M->np;
  D->bstore = xcalloc ((3 + 1 + 2 * (size_t) norder) * nl, sizeof (double));
  for (i = 0; i < 3; i++) {
    D->bcu[i] = D->bstore + i * nl;
    if (src[i]) memcpy (D->bcu[i], src[i], nl * sizeof (double));
  }
  D->bcp = D->bstore + 3 * nl;
  D->un = xcalloc (norder, sizeof (double*));
  D->hop = xcalloc (norder, sizeof (double*));
  for (q = 0; q < norder; q++) {
    D->un[q] = D->bstore + (4 + 2 * (size_t) q) * nl;
    D->hop[q] = D->bstore + (5 + 2 * (size_t) q) * nl;
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (m = 0; m < nmodes; m++) {
    const double bk = beta * m;
    orc_sys_destroy (D->psys[m]);
    D->psys[m] = orc_sys_create (Mp, 0.0, bk * bk, ORC_DIRECT);
  }
  return D;
}

int orc_dns_set_iterative (orc_dns* D, int level)
/* The ITERATIVE token (dns -i, dns/integrate.cpp:317-362 preSolve):
 * level >= 1 solves the velocity with JACPCG, level >= 2 the pressure with
 * MIXED (mode 0 DIRECT, the others JACPCG).  Before the first step.     */
{
  const int nmodes = (D->nz > 1) ? D->nz / 2 : 1;
  double alpha[4];
  int m;
  if (D->step || level < 0 || level > 2) return 1;
  stiffly_stable (D->nord, alpha);
  for (m = 0; m < nmodes; m++) {
    const double bk = D->beta * m;
    orc_sys_destroy (D->vsys[m]);
    orc_sys_destroy (D->psys[m]);
    D->vsys[m] = orc_sys_create (D->M, alpha[0] / (D->dt * D->kinvis), bk * bk,
                                 level >= 1 ? ORC_JACPCG : ORC_DIRECT);
    D->psys[m] = orc_sys_create (D->Mp, 0.0, bk * bk,
                                 (level >= 2 && m > 0) ? ORC_JACPCG : ORC_DIRECT);
  }
  return 0;
}

void orc_dns_set_bc (orc_dns* D, const double* bcu, const double* bcv, const double* bcw)
/* New velocity BC lines for the steps that follow (the re-evaluation of
 * time-dependent BCs, dns/integrate.cpp:196-201).                       */
{
  const double* src[3] = {bcu, bcv, bcw};
  const size_t nl = (size_t)D->nz * D->nline;
  int i;
  for (i = 0; i < 3; i++)
    if (src[i]) memcpy (D->bcu[i], src[i], nl * sizeof (double));
}

void orc_dns_destroy (orc_dns* D)
{
  free (D->bstore); free (D->un); free (D->hop);
  const int nmodes = (D->nz > 1) ? D->nz / 2 : 1;
  int m;
  for (m = 0; m < nmodes; m++) { orc_sys_destroy (D->vsys[m]); orc_sys_destroy (D->psys[m]); }
  free (D->vsys); free (D->psys); free (D->Us); free (D->Uf); free (D->store); free (D);
}

static void transform (const orc_dns* D, double* a, int sign)
/* AuxField::transform (src/auxfield.cpp:1133-1170), one process.        */
{
  const orc_mesh* M = D->M;
  if (D->nz == 1) return;
  if (D->nz == 2) {
    if (sign == 1) memset (a + M->planeSize, 0, M->planeSize * sizeof (double));
    else memcpy (a + M->planeSize, a, M->planeSize * sizeof (double));
    return;
  }
  orc_dftr (a, D->nz, M->planeSize, sign);
}

static void field_solve (orc_dns* D, orc_sys** sys, double lambda2_tmp,
                         int use_tmp, double* f, double* u, const double* bc)
/* bc: nz BC lines of the field's mesh, or NULL without boundaries       */
{
  const orc_mesh* M = sys ? sys[0]->M : D->M;
  const int nz = D->nz, nmodes = (nz > 1) ? nz / 2 : 1;
  orc_sys** s = sys;
  int k, m;
  if (use_tmp) {
    s = xcalloc (nmodes, sizeof (orc_sys*));
#pragma omp parallel for schedule(dynamic, 1)
    for (m = 0; m < nmodes; m++) {
      const double bk = D->beta * m;
      s[m] = orc_sys_create (M, lambda2_tmp, bk * bk, ORC_JACPCG);
    }
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (k = 0; k < nz; k++) {
    if (nz > 1 && k == 1) continue;
    orc_solve_plane (s[k >> 1], f + (size_t) k * M->planeSize,
                     bc ? bc + (size_t) k * M->nbound * M->np : NULL,
                     u + (size_t) k * M->planeSize, D->tol_rel, D->step_max);
  }
  if (use_tmp) {
    for (m = 0; m < nmodes; m++) orc_sys_destroy (s[m]);
    free (s);
  }
}

static void times (size_t n, double* t, const double* a, const double* b)
{ size_t i; for (i = 0; i < n; i++) t[i] = a[i] * b[i]; }
static void times_minus (size_t n, double* t, const double* a, const double* b)
{ size_t i; for (i = 0; i < n; i++) t[i] -= a[i] * b[i]; }

static void nonlinear (orc_dns* D, double** u, double* scratch)
/* skewSymmetric (dns/nonlinear.cpp:36-211, Cartesian :179-209),
 * altSkewSymmetric (:214-411, Cartesian branch), convective (:414-525,
 * Cartesian :502-524: u . grad u only, not halved) and Stokes (:528-570:
 * no nonlinear terms; advection 0-3 as dns -S / default / -C / -N,
 * dns/drive.cpp:114-158).  On exit u[] hold the physical-space
 * velocities, Us[0] the Fourier ones, Uf[0] N(u).                        */
{
  const orc_mesh* M = D->M;
  const int ncom = D->ncom, ndim = (D->nz > 1) ? 3 : 2;
  const size_t n = D->ntot;
  double* tmp = scratch;
  int i, j;
  for (i = 0; i < ncom; i++) {
    memset (D->Uf[i], 0, n * sizeof (double));
    memcpy (D->Us[i], u[i], n * sizeof (double));
    transform (D, u[i], -1);
  }
  for (i = 0; i < ncom; i++) {
    double* N = D->Uf[i];
    const int conv = (D->advection == 0) || (D->advection == 2) ||
                     (D->advection == 1 && D->toggle);
    const int cons = (D->advection == 0) || (D->advection == 1 && !D->toggle);
    for (j = 0; j < ndim; j++) {
      if (conv) {   /* n_i -= u_j d(u_i)/dx_j */
        if (j == 2) {
          memcpy (tmp, D->Us[i], n * sizeof (double));
          orc_gradient (M, D->nz, D->beta, tmp, 2);
          transform (D, tmp, -1);
        } else {
          memcpy (tmp, u[i], n * sizeof (double));
          orc_gradient (M, D->nz, D->beta, tmp, j);
        }
        times_minus (n, N, u[j], tmp);
      }
      if (cons) {   /* n_i -= d(u_i u_j)/dx_j */
        times (n, tmp, u[i], u[j]);
        if (j == 2) transform (D, tmp, +1);
        orc_gradient (M, D->nz, D->beta, tmp, j);
        if (j == 2) transform (D, tmp, -1);
        { size_t q; for (q = 0; q < n; q++) N[q] -= tmp[q]; }
      }
    }
    if (D->advection == 0) { size_t q; for (q = 0; q < n; q++) N[q] *= 0.5; }
    transform (D, N, +1);
    orc_smooth (M, D->nz, N);
  }
  if (D->advection == 1) D->toggle = 1 - D->toggle;
}

static void roll (double** lev, int nord, int ncom)
/* rollm (src/sem.h:106-120)                                             */
{
  int i, j;
  if (nord < 2) return;
  for (j = 0; j < ncom; j++) {
    double* t = lev[(nord-1)*ncom + j];
    for (i = nord - 1; i; i--) lev[i*ncom + j] = lev[(i-1)*ncom + j];
    lev[j] = t;
  }
}

int orc_dns_run (orc_dns* D, int nstep, double* u0, double* v0, double* w0,
                 double* p, double* t_step)
/* integrate (dns/integrate.cpp:79-221) with waveProp (:224-259),
 * setPForce (:262-277), project (:280-314), Solve (:365-395).           */
{
  const orc_mesh* M = D->M;
  const int ncom = D->ncom, nord = D->nord, ndim = (D->nz > 1) ? 3 : 2;
  const size_t n = D->ntot;
  double* u[3];
  double* scratch = xcalloc (n, sizeof (double));
  double alpha[4], beta[3];
  int s, i, q;
  double t0 = omp_get_wtime ();
  u[0] = u0; u[1] = v0; u[2] = w0;
  if (D->nthreads > 0) omp_set_num_threads (D->nthreads);

  for (s = 0; s < nstep; s++) {
    int Je;
    nonlinear (D, u, p);            /* pressure storage is scratch here  */
    D->step += 1;
    D->time += D->dt;
    /* high-order pressure BC storage, then the pressure BCs
     * (dns/integrate.cpp:168-172)                                       */
    if (D->bstore) {
      maintain_fourier (D, D->Us, D->Uf);
      evaluate_pressure_bcs (D);
    }
    Je = (D->step < nord) ? D->step : nord;
    stiffly_stable (Je, alpha);
    extrapolation (Je, beta);
    for (q = 0; q < Je; q++) beta[q] *= D->dt;
    /* waveProp */
    for (i = 0; i < ncom; i++) {
      size_t k;
      memset (u[i], 0, n * sizeof (double));
      for (q = 0; q < Je; q++) {
        const double* us = D->Us[q*ncom + i];
        const double* uf = D->Uf[q*ncom + i];
        for (k = 0; k < n; k++) u[i][k] += -alpha[q+1] * us[k];
        for (k = 0; k < n; k++) u[i][k] += beta[q] * uf[k];
      }
    }
    /* swapData (D->u[i], Us[0][i]) */
    for (i = 0; i < ncom; i++) {
      memcpy (scratch, u[i], n * sizeof (double));
      memcpy (u[i], D->Us[i], n * sizeof (double));
      memcpy (D->Us[i], scratch, n * sizeof (double));
    }
    roll (D->Uf, nord, ncom);
    /* setPForce */
    for (i = 0; i < ndim; i++) {
      memcpy (D->Uf[i], D->Us[i], n * sizeof (double));
      orc_gradient (M, D->nz, D->beta, D->Uf[i], i);
    }
    for (i = 1; i < ndim; i++) { size_t k; for (k = 0; k < n; k++) D->Uf[0][k] += D->Uf[i][k]; }
    { size_t k; for (k = 0; k < n; k++) D->Uf[0][k] /= D->dt; }
    /* pressure solve */
    field_solve (D, D->psys, 0.0, 0, D->Uf[0], p, D->bstore ? D->bcp : NULL);
    /* project */
    {
      const double a = -1.0 / (D->dt * D->kinvis), b = 1.0 / D->kinvis;
      for (i = 0; i < ncom; i++) {
        double* t = D->Us[i]; D->Us[i] = D->Uf[i]; D->Uf[i] = t;
        { size_t k; for (k = 0; k < n; k++) D->Uf[i][k] *= a; }
      }
      for (i = 0; i < ndim; i++) {
        size_t k;
        memcpy (D->Us[0], p, n * sizeof (double));
        orc_gradient (M, D->nz, D->beta, D->Us[0], i);
        for (k = 0; k < n; k++) D->Uf[i][k] += b * D->Us[0][k];
      }
    }
    for (i = 0; i < ncom; i++) memcpy (D->Us[i], u[i], n * sizeof (double));
    roll (D->Us, nord, ncom);
    /* viscous solves */
    for (i = 0; i < ncom; i++) {
      if (D->step < nord) {
        double a2[4];
        stiffly_stable (D->step, a2);
        field_solve (D, NULL, a2[0] / (D->dt * D->kinvis), 1, D->Uf[i], u[i],
                     D->bstore ? D->bcu[i] : NULL);
      } else
        field_solve (D, D->vsys, 0.0, 0, D->Uf[i], u[i], D->bstore ? D->bcu[i] : NULL);
    }
  }
  if (t_step) *t_step = (omp_get_wtime () - t0) / (nstep ? nstep : 1);
  free (scratch);
  return 0;
}
