// semx_dns.hip -- field operators and the 3D3C Navier-Stokes time step of
// semtex's dns solver on CDNA4 (gfx950):
//
//   k_dft_z      dDFTr along z, femlib/fourier.c:47-171: every column of a
//                field (ntrn columns, z-stride ntrn) transformed as one
//                dense real N x N product with semtex's packing
//                [Re0, Re_N/2, Re1, Im1, ...] and 1/N forward scaling
//   k_grad_xy    AuxField::gradient x/y, src/auxfield.cpp:514-665 with
//                Element::grad, src/element.cpp:552-597
//   k_grad_z     AuxField::gradient z in Fourier space, auxfield.cpp:636-657
//   k_smooth     Field::smooth, src/field.cpp:190-236 (mass-weighted DSS)
//   elementwise  AuxField times / timesMinus / axpy / scaling passes
//
//   k_hopbc      BCmgr::maintainFourier and evaluateCNBCp (src/bcmgr.cpp:
//                1039-1337): the computed high-order pressure flux on the
//                boundary sides of a bounded mesh
//
// and the host driver of integrate (dns/integrate.cpp:79-221) with
// waveProp, setPForce, project and the velocity/pressure solves of
// Field::solve (semx_solve.hip), on a periodic mesh or on a bounded one
// (velocity BC lines given, pressure flux computed).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "semx_hip.h"

using namespace semx;

namespace {

constexpr int DFT_KB = 16;          // padding of the transform matrix columns
constexpr int OP_THREADS = 1024;    // element-operator blocks

// The same product on the FP64 matrix cores (v_mfma_f64_16x16x4: A 16x4,
// B 4x16, D 16x16; lane l holds A[l%16][l/16], B[l/16][l%16] and, in
// register r, D[4r + l/16][l%16]).  W (MT 16-row tiles by K padded to 4) is
// staged in LDS with an odd row stride, so the 16 rows an A operand reads
// fall in distinct banks; each wave owns DFT_MCT 16-column tiles of the
// block's columns and all MT row tiles, and takes its B operands (four input
// rows by 16 columns per k step) straight from memory, DFT_MPF steps ahead
// in a register ring.  Input rows at or beyond N are clamped (their W
// columns are zero), as are columns beyond ntrn (not stored).  For N <= 128.
constexpr int DFT_MCT = 2;                       // column tiles per wave
constexpr int DFT_MPF = 4;                       // k steps of B prefetch
constexpr int DFT_MCOLS = 4 * 16 * DFT_MCT;      // columns per block
typedef double dft_d4 __attribute__((ext_vector_type(4)));

template <int MT>
__global__ __launch_bounds__(256) void k_dft_mfma(const double* __restrict__ in,
                                                  double* __restrict__ out,
                                                  const double* __restrict__ W, int N, int ldw,
                                                  int ntrn, long long ld) {
  extern __shared__ double sWm[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int KP = (N + 3) & ~3;
  const int ldk = KP + 1;
  const int li = lane & 15, lk = lane >> 4;
  int col[DFT_MCT];
#pragma unroll
  for (int ct = 0; ct < DFT_MCT; ++ct)
    col[ct] = min(blockIdx.x * DFT_MCOLS + (wave * DFT_MCT + ct) * 16 + li, ntrn - 1);
  // B operands DFT_MPF k steps ahead in a register ring (rows clamped)
  const int KS = KP / 4;
  auto load_b = [&](int ks, double (&b)[DFT_MCT]) {
    const size_t row = (size_t)min(4 * ks + lk, N - 1) * ld;
#pragma unroll
    for (int ct = 0; ct < DFT_MCT; ++ct) b[ct] = in[row + col[ct]];
  };
  double bq[DFT_MPF][DFT_MCT];
#pragma unroll
  for (int u = 0; u < DFT_MPF; ++u) load_b(u, bq[u]);
  for (int r = tid >> 6; r < MT * 16; r += 4)
    for (int k = lane; k < KP; k += 64) sWm[r * ldk + k] = W[(size_t)r * ldw + k];
  __syncthreads();
  dft_d4 acc[MT][DFT_MCT];
#pragma unroll
  for (int rt = 0; rt < MT; ++rt)
#pragma unroll
    for (int ct = 0; ct < DFT_MCT; ++ct) acc[rt][ct] = dft_d4{0.0, 0.0, 0.0, 0.0};
  const double* wa = sWm + li * ldk + lk;
  for (int k0 = 0; k0 < KS; k0 += DFT_MPF) {
#pragma unroll
    for (int u = 0; u < DFT_MPF; ++u) {
      const int ks = k0 + u;
      if (ks < KS) {
#pragma unroll
        for (int rt = 0; rt < MT; ++rt) {
          const double a = wa[rt * 16 * ldk + 4 * ks];
#pragma unroll
          for (int ct = 0; ct < DFT_MCT; ++ct)
            acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bq[u][ct], acc[rt][ct], 0, 0, 0);
        }
      }
      load_b(ks + DFT_MPF, bq[u]);
    }
  }
#pragma unroll
  for (int ct = 0; ct < DFT_MCT; ++ct) {
    const int c = blockIdx.x * DFT_MCOLS + (wave * DFT_MCT + ct) * 16 + li;
    if (c >= ntrn) continue;
#pragma unroll
    for (int rt = 0; rt < MT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + 4 * r + lk;
        if (row < N) out[(size_t)row * ld + c] = acc[rt][ct][r];
      }
  }
}

// out[r][col] = sum_k W[r][k] in[k][col], r < N: a register-blocked
// product of the (padded) N x N transform matrix with a 64-column tile of
// the field.  Threads form 16 row groups x 16 column groups; each computes
// RP rows x 4 columns, with the k dimension staged through LDS 16 rows at a
// time.  W has rows padded to 16*RP and columns to a multiple of 16 (zero
// entries), so loads of rows >= N are clamped without a branch.
constexpr int DFT_COLS = 64;
constexpr int DFT_KC = 16;

template <int RP>
__global__ __launch_bounds__(256) void k_dft_z(const double* __restrict__ in,
                                               double* __restrict__ out,
                                               const double* __restrict__ W, int N, int ldw,
                                               int ncol, long long stride) {
  __shared__ double sX[DFT_KC][DFT_COLS];
  __shared__ double sW[16 * RP][DFT_KC + 1];
  const int tid = threadIdx.x;
  const int cg = tid & 15, rg = tid >> 4;
  const int col0 = blockIdx.x * DFT_COLS;
  double acc[RP][4];
#pragma unroll
  for (int i = 0; i < RP; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < N; k0 += DFT_KC) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < DFT_KC * DFT_COLS / 256; ++q) {
      const int idx = tid + q * 256;
      const int kk = idx / DFT_COLS, c = idx - (idx / DFT_COLS) * DFT_COLS;
      sX[kk][c] = in[(size_t)min(k0 + kk, N - 1) * stride + min(col0 + c, ncol - 1)];
    }
#pragma unroll
    for (int q = 0; q < RP; ++q) {
      const int idx = tid + q * 256;            // 16*RP rows x 16 k
      const int r = idx / DFT_KC, kk = idx - (idx / DFT_KC) * DFT_KC;
      sW[r][kk] = W[(size_t)r * ldw + k0 + kk];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < DFT_KC; ++kk) {
      double xv[4], wv[RP];
#pragma unroll
      for (int j = 0; j < 4; ++j) xv[j] = sX[kk][cg * 4 + j];
#pragma unroll
      for (int i = 0; i < RP; ++i) wv[i] = sW[rg * RP + i][kk];
#pragma unroll
      for (int i = 0; i < RP; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += wv[i] * xv[j];
    }
  }
#pragma unroll
  for (int i = 0; i < RP; ++i) {
    const int r = rg * RP + i;
    if (r >= N) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = col0 + cg * 4 + j;
      if (c < ncol) out[(size_t)r * stride + c] = acc[i][j];
    }
  }
}

// d/dx (dir 0) or d/dy (dir 1) on every element of every plane:
// tgt = (D_s u) * ds/dx_i + (u D_r^T) * dr/dx_i, plain (non-SVV) D.
__global__ __launch_bounds__(OP_THREADS) void k_grad_xy(double* __restrict__ data,
                                                        const double* __restrict__ G,
                                                        const int* __restrict__ eset,
                                                        const double* __restrict__ ops, int np,
                                                        int nel, int planeSize, int dir) {
  extern __shared__ double gsh[];
  const int npnp = np * np;
  const int ce = OP_THREADS / npnp;
  const int t = threadIdx.x;
  const int el = t / npnp, pt = t - el * npnp;
  const int e = blockIdx.x * ce + el;
  const bool act = el < ce && e < nel;
  double* sD = gsh;
  double* sT = sD + npnp;
  double* sU = sT + npnp;
  if (t < npnp) {
    sD[t] = ops[t];
    sT[t] = ops[npnp + t];
  }
  double* u = data + (size_t)blockIdx.y * planeSize + (size_t)e * npnp;
  if (act) sU[el * npnp + pt] = u[pt];
  __syncthreads();
  if (!act) return;
  const int i = pt / np, j = pt - (pt / np) * np;
  const double* su = sU + el * npnp;
  double A = 0.0, B = 0.0;
  for (int q = 0; q < np; ++q) {
    A += su[i * np + q] * sT[q * np + j];
    B += sD[i * np + q] * su[q * np + j];
  }
  const double* g = G + (size_t)eset[e] * G_N * npnp;
  const double rx = g[(dir ? G_DRDY : G_DRDX) * npnp + pt];
  const double sx = g[(dir ? G_DSDY : G_DSDX) * npnp + pt];
  u[pt] = B * sx + A * rx;
}

// An x or y derivative (the element operator of k_grad_xy) read out of
// place, so the operand needs no copy, with its use applied in the
// epilogue:
//   MODE 0  N -= b * d(a)/dx_dir     convective term (dns/nonlinear.cpp:36-211)
//   MODE 1  N -= d(a b)/dx_dir       conservative term
//   MODE 2  N  = d(a)/dx_dir         setPForce (dns/integrate.cpp:159-160)
//   MODE 3  N = sc N + c d(a)/dx_dir project (dns/integrate.cpp:174-186)
// Entries of a plane beyond its elements (the padding up to planeSize) take
// the values the unfused copy-derivative-update sequence gives them (the
// derivative leaves them as they were): N -= a b, N -= a b, N = a,
// N = sc N + c a.  In modes 0 and 1, sc = 0 starts N from zero (N is not
// read: the first term of a nonlinear product needs no cleared field).
template <int MODE>
__global__ __launch_bounds__(OP_THREADS) void k_grad_nl(double* __restrict__ N,
                                                        const double* __restrict__ a,
                                                        const double* __restrict__ b,
                                                        const double* __restrict__ G,
                                                        const int* __restrict__ eset,
                                                        const double* __restrict__ ops, int np,
                                                        int nel, int planeSize, int dir,
                                                        double c, double sc) {
  extern __shared__ double gsh[];
  const int npnp = np * np;
  const int ce = OP_THREADS / npnp;
  const int t = threadIdx.x;
  const int el = t / npnp, pt = t - el * npnp;
  const int e = blockIdx.x * ce + el;
  const bool act = el < ce && e < nel;
  double* sD = gsh;
  double* sT = sD + npnp;
  double* sU = sT + npnp;
  if (t < npnp) {
    sD[t] = ops[t];
    sT[t] = ops[npnp + t];
  }
  const size_t pbase = (size_t)blockIdx.y * planeSize;
  if (blockIdx.x == 0)
    for (int k = nel * npnp + t; k < planeSize; k += OP_THREADS) {
      const size_t q = pbase + k;
      if (MODE <= 1) N[q] = (sc == 0.0 ? 0.0 : N[q]) - a[q] * b[q];
      else if (MODE == 2) N[q] = a[q];
      else N[q] = sc * N[q] + c * a[q];
    }
  const size_t off = pbase + (size_t)e * npnp + pt;
  if (act) sU[el * npnp + pt] = MODE == 1 ? a[off] * b[off] : a[off];
  __syncthreads();
  if (!act) return;
  const int i = pt / np, j = pt - (pt / np) * np;
  const double* su = sU + el * npnp;
  double A = 0.0, B = 0.0;
  for (int q = 0; q < np; ++q) {
    A += su[i * np + q] * sT[q * np + j];
    B += sD[i * np + q] * su[q * np + j];
  }
  const double* g = G + (size_t)eset[e] * G_N * npnp;
  const double rx = g[(dir ? G_DRDY : G_DRDX) * npnp + pt];
  const double sx = g[(dir ? G_DSDY : G_DSDX) * npnp + pt];
  const double d = B * sx + A * rx;
  if (MODE == 0) N[off] = (sc == 0.0 ? 0.0 : N[off]) - b[off] * d;
  else if (MODE == 1) N[off] = (sc == 0.0 ? 0.0 : N[off]) - d;
  else if (MODE == 2) N[off] = d;
  else N[off] = sc * N[off] + c * d;
}

// Computed pressure boundary conditions (KIO91; BCmgr::maintainFourier,
// src/bcmgr.cpp:1039-1298, and evaluateCNBCp, :1301-1337), Cartesian.  One
// workgroup per (boundary side, Fourier mode); at each node of the side and
// each plane of the mode (the Nyquist plane is left alone):
//   un   = u.n
//   hop  = -(1/dt) (a0 un + sum_{l<Jd} a_{l+1} un_l)      (Jd > 0 only)
//          + n.N(u) - nu n.curlCurl(u)
//   un_l, hop_l shift down one level (level 0 the new value)
//   bcp  = sum_{l<Je} b_l hop_l on the pressure's natural sides, else 0
// with u the Fourier velocity at the start of the step and N(u) its
// nonlinear terms.  curlCurl (Edge::curlCurl, src/edge.cpp:98-258) takes
// the z vorticity w = dv/dx - du/dy on the whole element and its gradient
// at the side: (dw/dy, -dw/dx), plus for mode k > 0 (beta k) the
// semi-Fourier terms of w-velocity and (beta k)^2 (u, v).  Each thread owns
// one (plane, side node) entry of every level array, so the in-place level
// shift has no hazards, and the level arrays never move (one HIP graph of
// the step stays valid).
struct HopArgs {
  const double* U[3];     // Fourier velocity, nz*planeSize each (W unused in 2D)
  const double* N[2];     // Fourier nonlinear terms, x and y
  double* un;             // nord levels of u.n, nz*nline each
  double* hop;            // nord levels of the flux terms
  double* bcp;            // pressure BC lines, nz*nline
  const int* bel;         // nbound: element
  const int* bsd;         // nbound: side 0-3
  const int* pnat;        // nbound: pressure side natural
  const double* nrm;      // 2*nline: unit outward normal x, then y
  const double* G;
  const int* eset;
  const double* ops;      // DV first
  int np, nz, nline, planeSize, nord, w3, Jd, Je;
  double beta, nu, invdt;
  double a[4], b[3];
};

__device__ __forceinline__ int side_node(int np, int s, int q) {
  return s == 0 ? q : s == 1 ? q * np + np - 1 : s == 2 ? (np - 1) * np + (np - 1 - q)
                                                        : (np - 1 - q) * np;
}

// x and y derivatives at element node n of element data f (LDS, np x np).
__device__ __forceinline__ void node_grad(const double* f, const double* sD, int np, int n,
                                          const double* g, int npnp, double& gx, double& gy) {
  const int i = n / np, j = n - (n / np) * np;
  double ddr = 0.0, dds = 0.0;
  for (int k = 0; k < np; ++k) {
    ddr += sD[j * np + k] * f[i * np + k];
    dds += sD[i * np + k] * f[k * np + j];
  }
  gx = ddr * g[G_DRDX * npnp + n] + dds * g[G_DSDX * npnp + n];
  gy = ddr * g[G_DRDY * npnp + n] + dds * g[G_DSDY * npnp + n];
}

__global__ __launch_bounds__(256) void k_hopbc(HopArgs A) {
  __shared__ double sD[256], sf[6][256], sw[2][256];
  const int np = A.np, npnp = np * np, t = threadIdx.x;
  const int b = blockIdx.x, m = blockIdx.y;
  const int e = A.bel[b], sd = A.bsd[b];
  const int npl = (m == 0) ? 1 : 2;               // planes of this mode
  const int k0 = 2 * m;                           // mode 0: plane 0 only
  const bool zt = m > 0 && A.w3;                  // semi-Fourier terms of w
  const double* g = A.G + (size_t)A.eset[e] * G_N * npnp;
  if (t < npnp) {
    sD[t] = A.ops[t];
    for (int p = 0; p < npl; ++p) {
      const size_t off = (size_t)(k0 + p) * A.planeSize + (size_t)e * npnp + t;
      sf[2 * p][t] = A.U[0][off];
      sf[2 * p + 1][t] = A.U[1][off];
      if (zt) sf[4 + p][t] = A.U[2][off];
    }
  }
  __syncthreads();
  // z vorticity of each plane on the whole element
  if (t < npnp)
    for (int p = 0; p < npl; ++p) {
      double ux, uy, vx, vy;
      node_grad(sf[2 * p], sD, np, t, g, npnp, ux, uy);
      node_grad(sf[2 * p + 1], sD, np, t, g, npnp, vx, vy);
      sw[p][t] = vx - uy;
    }
  __syncthreads();
  if (t >= npl * np) return;
  const int p = t / np, q = t - (t / np) * np;    // p 0 real, 1 imaginary
  const int n = side_node(np, sd, q);
  const double bk = A.beta * m;
  double wx, wy;
  node_grad(sw[p], sD, np, n, g, npnp, wx, wy);
  double cx = wy, cy = -wx;                       // curl curl (u, v) of the plane
  if (zt) {   // real: -bk d(Wi)/dx_j, imaginary: +bk d(Wr)/dx_j
    double zx, zy;
    node_grad(sf[4 + (1 - p)], sD, np, n, g, npnp, zx, zy);
    const double sg = p ? bk : -bk;
    cx += sg * zx;
    cy += sg * zy;
  }
  if (m > 0) {
    cx += bk * bk * sf[2 * p][n];
    cy += bk * bk * sf[2 * p + 1][n];
  }
  const int j = b * np + q;
  const size_t i = (size_t)(k0 + p) * A.nline + j;
  const size_t lv = (size_t)A.nz * A.nline;
  const double nx = A.nrm[j], ny = A.nrm[A.nline + j];
  const size_t off = (size_t)(k0 + p) * A.planeSize + (size_t)e * npnp + n;
  const double un = sf[2 * p][n] * nx + sf[2 * p + 1][n] * ny;
  double h = 0.0;
  if (A.Jd) {
    double wk = A.a[0] * un;
    for (int l = 0; l < A.Jd; ++l) wk += A.a[l + 1] * A.un[l * lv + i];
    h = -A.invdt * wk;
  }
  h += A.N[0][off] * nx + A.N[1][off] * ny;
  h += -A.nu * cx * nx + -A.nu * cy * ny;
  for (int l = A.nord - 1; l > 0; --l) {
    A.un[l * lv + i] = A.un[(l - 1) * lv + i];
    A.hop[l * lv + i] = A.hop[(l - 1) * lv + i];
  }
  A.un[i] = un;
  A.hop[i] = h;
  double v = 0.0;
  if (A.pnat[b])
    for (int l = 0; l < A.Je; ++l) v += A.b[l] * A.hop[l * lv + i];
  A.bcp[i] = v;
}

// z-derivative of Fourier planes: mode k Re <- -beta k Im, Im <- beta k Re;
// mode 0 and the Nyquist plane become zero.  src may equal dst.
__global__ void k_grad_z(const double* src, double* dst, int nmodes, int planeSize, double beta) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (idx >= (size_t)planeSize || k >= nmodes) return;
  const size_t re = (size_t)(2 * k) * planeSize + idx, im = re + planeSize;
  if (k == 0) {
    dst[re] = 0.0;
    dst[im] = 0.0;
  } else {
    const double bk = beta * k;
    const double r = src[re], m = src[im];
    dst[re] = -bk * m;
    dst[im] = bk * r;
  }
}

// dst = sc dst + c d(src)/dz over the Fourier planes (k_grad_z's derivative).
__global__ void k_grad_z_axpby(const double* __restrict__ src, double* __restrict__ dst,
                               int nmodes, int planeSize, double beta, double sc, double c) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (idx >= (size_t)planeSize || k >= nmodes) return;
  const size_t re = (size_t)(2 * k) * planeSize + idx, im = re + planeSize;
  const double bk = beta * k;
  const double dr = k == 0 ? 0.0 : -bk * src[im], di = k == 0 ? 0.0 : bk * src[re];
  dst[re] = sc * dst[re] + c * dr;
  dst[im] = sc * dst[im] + c * di;
}

// Mass-weighted average of element-boundary values over the copies of each
// global node, per plane; copies are summed in increasing storage order.
__global__ void k_smooth(double* __restrict__ data, const double* __restrict__ G,
                         const int* __restrict__ eset, const int* __restrict__ emap,
                         const int* __restrict__ asm_ptr, const int* __restrict__ asm_idx,
                         const double* __restrict__ imass, int next, int npnp, int nglobal,
                         int planeSize) {
  const int gI = blockIdx.x * blockDim.x + threadIdx.x;
  if (gI >= nglobal) return;
  double* src = data + (size_t)blockIdx.y * planeSize;
  double s = 0.0;
  for (int c = asm_ptr[gI]; c < asm_ptr[gI + 1]; ++c) {
    const int e = asm_idx[c] / next, b = asm_idx[c] - (asm_idx[c] / next) * next;
    const int q = emap[b];
    s += G[((size_t)eset[e] * G_N + G_Q4) * npnp + q] * src[(size_t)e * npnp + q];
  }
  s *= imass[gI];
  for (int c = asm_ptr[gI]; c < asm_ptr[gI + 1]; ++c) {
    const int e = asm_idx[c] / next, b = asm_idx[c] - (asm_idx[c] / next) * next;
    src[(size_t)e * npnp + emap[b]] = s;
  }
}

// --- elementwise passes over whole fields -------------------------------
__global__ void k_times(size_t n, double* __restrict__ t, const double* __restrict__ a,
                        const double* __restrict__ b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = a[i] * b[i];
}

__global__ void k_times_minus(size_t n, double* __restrict__ t, const double* __restrict__ a,
                              const double* __restrict__ b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] -= a[i] * b[i];
}

__global__ void k_fill(size_t n, double* __restrict__ t, double a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = a;
}

__global__ void k_copy(size_t n, double* __restrict__ t, const double* __restrict__ a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = a[i];
}

__global__ void k_scale(size_t n, double* __restrict__ t, double a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] *= a;
}

// AuxField::couple / Field::coupleBCs (src/auxfield.cpp:1352-1420,
// src/field.cpp:1043-1118): per mode k >= klo, Re/Im planes 2k, 2k+1 of n
// values.  dir > 0: v~ = v + i w, w~ = v - i w; dir < 0: v = (v~ + w~) / 2,
// w = i (w~ - v~) / 2 -- the same sums as semtex's Veclib sequence.
__global__ void k_couple(double* __restrict__ v, double* __restrict__ w, size_t n, int klo,
                         int nmode, int dir) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = klo + (int)blockIdx.y;
  if (i >= n || k >= nmode) return;
  double* vr = v + (size_t)(2 * k) * n;
  double* vi = vr + n;
  double* wr = w + (size_t)(2 * k) * n;
  double* wi = wr + n;
  const double a = vr[i], b = vi[i], c = wr[i], d = wi[i];
  if (dir > 0) {
    vr[i] = a - d;
    vi[i] = b + c;
    wr[i] = d + a;
    wi[i] = b - c;
  } else {
    vr[i] = 0.5 * (a + c);
    vi[i] = 0.5 * (b + d);
    wr[i] = 0.5 * (b - d);
    wi[i] = 0.5 * (c - a);
  }
}

__global__ void k_axpy(size_t n, double* __restrict__ t, double a, const double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] += a * x[i];
}

__global__ void k_axpby(size_t n, double* __restrict__ t, double b, double a,
                        const double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = b * t[i] + a * x[i];
}

struct WaveArgs {
  const double* us[3];
  const double* uf[3];
  double a[3], b[3];
  int nlev;
};

// waveProp (dns/integrate.cpp:224-259): u = sum_q (-alpha_{q+1} Us_q) +
// (beta_q dt Uf_q), accumulated in semtex's order.
__global__ void k_wave(size_t n, double* __restrict__ u, WaveArgs w) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int q = 0; q < w.nlev; ++q) {
    s += w.a[q] * w.us[q][i];
    s += w.b[q] * w.uf[q][i];
  }
  u[i] = s;
}

// setPForce (dns/integrate.cpp:262-277): f0 = (f0 + f1 [+ f2]) / dt
__global__ void k_divergence_sum(size_t n, double* __restrict__ f0, const double* __restrict__ f1,
                                 const double* __restrict__ f2, double dt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = f0[i] + f1[i];
  if (f2) s += f2[i];
  f0[i] = s / dt;
}

inline dim3 grid1(size_t n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// --- dense DFT matrices ----------------------------------------------------
bool fft_length_ok(int n) {
  if (n < 1) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

// rows per thread of k_dft_z for a transform of length N (16 row groups)
int dft_rp(int N) {
  const int rp = (N + 15) / 16;
  return rp <= 1 ? 1 : rp <= 2 ? 2 : rp <= 4 ? 4 : rp <= 6 ? 6 : rp <= 8 ? 8
         : rp <= 12 ? 12 : rp <= 16 ? 16 : -1;
}

// Row-major matrix of the real transform of length N (rows padded to
// 16 * dft_rp(N), columns to a multiple of DFT_KB, zero entries).  sign +1:
// semtex's forward packing with 1/N; sign -1: the unscaled inverse.  Angles
// are reduced exactly (jk mod N) and evaluated in long double.
std::vector<double> dft_matrix(int N, int sign, int* ldw) {
  const int rows = 16 * dft_rp(N);
  const int ld = (N + DFT_KB - 1) / DFT_KB * DFT_KB;
  *ldw = ld;
  std::vector<double> W((size_t)rows * ld, 0.0);
  const long double two_pi = 6.283185307179586476925286766559L;
  auto cs = [&](long long jk, long double* c, long double* s) {
    const long double a = two_pi * (long double)(jk % N) / (long double)N;
    *c = cosl(a);
    *s = sinl(a);
  };
  const int nh = N / 2;
  if (sign == 1) {
    for (int k = 0; k < N; ++k) {
      W[0 * ld + k] = (double)(1.0L / N);
      if (N > 1) W[1 * ld + k] = (double)(((k & 1) ? -1.0L : 1.0L) / N);
      for (int j = 1; j < nh; ++j) {
        long double c, s;
        cs((long long)j * k, &c, &s);
        W[(size_t)(2 * j) * ld + k] = (double)(c / N);
        W[(size_t)(2 * j + 1) * ld + k] = (double)(-s / N);
      }
    }
  } else {
    for (int k = 0; k < N; ++k) {
      double* r = &W[(size_t)k * ld];
      r[0] = 1.0;
      if (N > 1) r[1] = (k & 1) ? -1.0 : 1.0;
      for (int j = 1; j < nh; ++j) {
        long double c, s;
        cs((long long)j * k, &c, &s);
        r[2 * j] = (double)(2.0L * c);
        r[2 * j + 1] = (double)(-2.0L * s);
      }
    }
  }
  return W;
}

// The matrix of dft_matrix(N, sign) formed on the device, one thread per
// entry of its rows x ldw storage: the same packing, padding and scaling,
// with the angle reduced exactly (jk mod N) and evaluated by sincospi in
// FP64 (within a few units in the last place of the long-double values).
__global__ void k_dft_matrix(double* __restrict__ W, int N, int sign, int rows, int ldw) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= rows * ldw) return;
  const int r = idx / ldw, c = idx - (idx / ldw) * ldw;
  const int nh = N / 2;
  double v = 0.0;
  if (r < N && c < N) {
    // sign +1: row r is output coefficient r, column c input point c;
    // sign -1: row r is output point r, column c packed coefficient c
    const int pk = (sign == 1) ? r : c, pt = (sign == 1) ? c : r;
    const double sc = (sign == 1) ? 1.0 / N : 2.0;
    if (pk == 0) {
      v = (sign == 1) ? 1.0 / N : 1.0;
    } else if (pk == 1) {
      v = ((pt & 1) ? -1.0 : 1.0) * ((sign == 1) ? 1.0 / N : 1.0);
    } else if (pk / 2 < nh) {
      const long long m = ((long long)(pk / 2) * pt) % N;
      double sn, cn;
      sincospi(2.0 * (double)m / (double)N, &sn, &cn);
      v = (pk & 1) ? -sn : cn;
      v = (sign == 1) ? v / N : sc * v;
    }
  }
  W[idx] = v;
}

struct DftPlan { int N, sign, ldw; double* W; };
std::vector<DftPlan> g_plans;   // device matrices, kept for the process

int dft_plan(int N, int sign, const double** W, int* ldw) {
  for (const DftPlan& p : g_plans)
    if (p.N == N && p.sign == sign) { *W = p.W; *ldw = p.ldw; return SEMX_OK; }
  int ld = 0;
  std::vector<double> h = dft_matrix(N, sign, &ld);
  double* d = nullptr;
  int st = dev_upload(&d, h.data(), h.size());
  if (st) return st;
  g_plans.push_back({N, sign, ld, d});
  *W = d;
  *ldw = ld;
  return SEMX_OK;
}

// out-of-place transform of ntrn columns of length N (z-stride ntrn) with
// the transform matrix W (leading dimension ldw) already on the device
int dft_launch_w(const double* in, double* out, const double* W, int ldw, int N, int ntrn,
                 hipStream_t s) {
  if (N <= 128 && ntrn > 0) {
    // matrix cores: W staged whole in LDS
    const int MT = (N + 15) / 16;
    const size_t lds = (size_t)MT * 16 * (((N + 3) & ~3) + 1) * sizeof(double);
    dim3 gm((ntrn + DFT_MCOLS - 1) / DFT_MCOLS);
    const void* kfn = nullptr;
    switch (MT) {
      case 1: kfn = (const void*)k_dft_mfma<1>; break;
      case 2: kfn = (const void*)k_dft_mfma<2>; break;
      case 3: kfn = (const void*)k_dft_mfma<3>; break;
      case 4: kfn = (const void*)k_dft_mfma<4>; break;
      case 5: kfn = (const void*)k_dft_mfma<5>; break;
      case 6: kfn = (const void*)k_dft_mfma<6>; break;
      case 7: kfn = (const void*)k_dft_mfma<7>; break;
      default: kfn = (const void*)k_dft_mfma<8>; break;
    }
    if (lds > 64 * 1024)
      HIPCHK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    long long ldl = ntrn;
    void* kargs[] = {(void*)&in, (void*)&out, (void*)&W, (void*)&N, (void*)&ldw, (void*)&ntrn,
                     (void*)&ldl};
    hipEvent_t evm;
    prof_begin(s, &evm);
    HIPCHK(hipLaunchKernel(kfn, gm, dim3(256), kargs, lds, s));
    prof_end(s, "k_dft_z", evm);
    HIPCHK(hipGetLastError());
    return SEMX_OK;
  }
  const int rp = dft_rp(N);
  dim3 g((ntrn + DFT_COLS - 1) / DFT_COLS);
  hipEvent_t ev;
  prof_begin(s, &ev);
  switch (rp) {
    case 1: hipLaunchKernelGGL(k_dft_z<1>, g, dim3(256), 0, s, in, out, W, N, ldw, ntrn, (long long)ntrn); break;
    case 2: hipLaunchKernelGGL(k_dft_z<2>, g, dim3(256), 0, s, in, out, W, N, ldw, ntrn, (long long)ntrn); break;
    case 4: hipLaunchKernelGGL(k_dft_z<4>, g, dim3(256), 0, s, in, out, W, N, ldw, ntrn, (long long)ntrn); break;
    case 6: hipLaunchKernelGGL(k_dft_z<6>, g, dim3(256), 0, s, in, out, W, N, ldw, ntrn, (long long)ntrn); break;
    case 8: hipLaunchKernelGGL(k_dft_z<8>, g, dim3(256), 0, s, in, out, W, N, ldw, ntrn, (long long)ntrn); break;
    case 12: hipLaunchKernelGGL(k_dft_z<12>, g, dim3(256), 0, s, in, out, W, N, ldw, ntrn, (long long)ntrn); break;
    case 16: hipLaunchKernelGGL(k_dft_z<16>, g, dim3(256), 0, s, in, out, W, N, ldw, ntrn, (long long)ntrn); break;
    default: return SEMX_E_UNSUPPORTED;
  }
  prof_end(s, "k_dft_z", ev);
  HIPCHK(hipGetLastError());
  return SEMX_OK;
}

// the same with the process-wide plan of (N, sign)
int dft_launch(const double* in, double* out, int N, int ntrn, int sign, hipStream_t s) {
  const double* W = nullptr;
  int ldw = 0;
  int st = dft_plan(N, sign, &W, &ldw);
  if (st) return st;
  return dft_launch_w(in, out, W, ldw, N, ntrn, s);
}

// The fused derivative of k_grad_nl (modes 0-3 there), dir 0 or 1
int grad_nl_launch(Mesh& M, int nz, double* N, const double* a, const double* b, int dir,
                   int mode, hipStream_t s, double c = 0.0, double sc = 1.0) {
  hipEvent_t ev;
  const int ce = OP_THREADS / M.npnp;
  dim3 g((M.nel + ce - 1) / ce, nz);
  const size_t lds = (2 + (size_t)ce) * M.npnp * sizeof(double);
  prof_begin(s, &ev);
#define SEMX_GNL(MODE_)                                                                     \
  hipLaunchKernelGGL(k_grad_nl<MODE_>, g, dim3(OP_THREADS), lds, s, N, a, b, M.d.G, M.d.eset, \
                     M.d.ops, M.np, M.nel, M.planeSize, dir, c, sc)
  switch (mode) {
    case 0: SEMX_GNL(0); break;
    case 1: SEMX_GNL(1); break;
    case 2: SEMX_GNL(2); break;
    default: SEMX_GNL(3); break;
  }
#undef SEMX_GNL
  prof_end(s, "k_grad_nl", ev);
  HIPCHK(hipGetLastError());
  return SEMX_OK;
}

int grad_launch(Mesh& M, int nz, double beta, double* data, int dir, hipStream_t s) {
  hipEvent_t ev;
  if (dir < 2) {
    const int ce = OP_THREADS / M.npnp;
    dim3 g((M.nel + ce - 1) / ce, nz);
    const size_t lds = (2 + (size_t)ce) * M.npnp * sizeof(double);
    prof_begin(s, &ev);
    hipLaunchKernelGGL(k_grad_xy, g, dim3(OP_THREADS), lds, s, data, M.d.G, M.d.eset, M.d.ops,
                       M.np, M.nel, M.planeSize, dir);
    prof_end(s, "k_grad_xy", ev);
  } else {
    dim3 g((M.planeSize + 255) / 256, nz / 2);
    prof_begin(s, &ev);
    hipLaunchKernelGGL(k_grad_z, g, dim3(256), 0, s, data, data, nz / 2, M.planeSize, beta);
    prof_end(s, "k_grad_z", ev);
  }
  HIPCHK(hipGetLastError());
  return SEMX_OK;
}

int smooth_launch(Mesh& M, int nz, double* data, hipStream_t s) {
  hipEvent_t ev;
  dim3 g((M.nglobal + 255) / 256, nz);
  prof_begin(s, &ev);
  hipLaunchKernelGGL(k_smooth, g, dim3(256), 0, s, data, M.d.G, M.d.eset, M.d.emap, M.d.asm_ptr,
                     M.d.asm_idx, M.d.imass, M.next, M.npnp, M.nglobal, M.planeSize);
  prof_end(s, "k_smooth", ev);
  HIPCHK(hipGetLastError());
  return SEMX_OK;
}

// Integration coefficients, src/integration.cpp:99-164.
void stiffly_stable(int n, double* c) {
  if (n == 1) { c[0] = 1.0; c[1] = -1.0; }
  else if (n == 2) { c[0] = 1.5; c[1] = -2.0; c[2] = 0.5; }
  else { c[0] = 11.0 / 6.0; c[1] = -3.0; c[2] = 1.5; c[3] = -1.0 / 3.0; }
}

void extrapolation(int n, double* c) {
  if (n == 1) { c[0] = 1.0; }
  else if (n == 2) { c[0] = 2.0; c[1] = -1.0; }
  else { c[0] = 3.0; c[1] = -3.0; c[2] = 1.0; }
}

}  // namespace

// ===========================================================================
// DNS state: the fields of integrate() held on the device.
// ===========================================================================

// The buffer roles of the DNS state: which allocation each field of the
// step uses.  The step permutes them (transform output <-> spare,
// swapData, the rolled time levels) and returns to a given assignment only
// after a full cycle of steps (88 steps for N_TIME = 3 with the
// skew-symmetric terms, 182 with the alternating form).
struct DnsRoles {
  double* U[3];
  double* P;
  double* Us[3][3];
  double* Uf[3][3];
  double* tmp;
  double* spare;
  int toggle;
  bool operator==(const DnsRoles& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

// One steady step captured for one assignment of roles: replaying it leaves
// the roles as `after`.
struct DnsGraph {
  DnsRoles before, after;
  hipGraphExec_t exec;
};

struct semx_dns_s {
  Mesh* M = nullptr;
  int nz = 1, ncom = 2, ndim = 2, nord = 2, advection = 0, step = 0, step_max = 500, toggle = 1;
  double beta = 1.0, dt = 0.0, kinvis = 0.0, tol_rel = 1e-12;
  size_t ntot = 0;
  double* store = nullptr;              // all fields in one allocation
  double* U[3] = {nullptr, nullptr, nullptr};
  double* P = nullptr;
  double* Us[3][3] = {};                // [level][component]
  double* Uf[3][3] = {};
  double* tmp = nullptr;
  double* spare = nullptr;              // transform output, swapped in
  Helm* vel = nullptr;                  // DIRECT, lambda2 = alpha0/(dt nu)
  Helm* pre = nullptr;                  // DIRECT, lambda2 = 0
  Helm* start[2] = {nullptr, nullptr};  // JACPCG start-up velocity systems
  // bounded mesh: the pressure's numbering and side kinds (same elements
  // and boundary list as the velocity mesh), BC lines and flux history
  Mesh* Mp = nullptr;
  int nline = 0;
  double* bstore = nullptr;             // the arrays below in one allocation
  double* bcu[3] = {nullptr, nullptr, nullptr};   // velocity BC lines, nz*nline each
  double* bcp = nullptr;                // pressure BC lines
  double* un = nullptr;                 // nord levels of u.n
  double* hop = nullptr;                // nord levels of the flux terms
  double* nrm = nullptr;                // side normals, 2*nline
  int* bint = nullptr;                  // element, side, pressure-natural per boundary
  bool graphs_on = false;               // replay steady steps from HIP graphs
  hipStream_t capture = nullptr;        // stream the steps are captured on
  std::vector<DnsGraph> graphs;         // one per role assignment met
  long long replays = 0;
};

namespace {

int dns_transform(semx_dns_s& D, double*& a, int sign, hipStream_t s) {
  Mesh& M = *D.M;
  if (D.nz == 1) return SEMX_OK;
  if (D.nz == 2) {   // AuxField::transform with two planes (auxfield.cpp:1133-1170)
    const size_t n = M.planeSize;
    if (sign == 1) hipLaunchKernelGGL(k_fill, grid1(n, 256), dim3(256), 0, s, n, a + n, 0.0);
    else hipLaunchKernelGGL(k_copy, grid1(n, 256), dim3(256), 0, s, n, a + n, a);
    HIPCHK(hipGetLastError());
    return SEMX_OK;
  }
  int st = dft_launch(a, D.spare, D.nz, M.planeSize, sign, s);
  if (st) return st;
  std::swap(a, D.spare);
  return SEMX_OK;
}

int ew_launch_check() {
  HIPCHK(hipGetLastError());
  return SEMX_OK;
}

#define EW(kernel, ...)                                                          \
  do {                                                                           \
    hipLaunchKernelGGL(kernel, grid1(n, 256), dim3(256), 0, s, n, __VA_ARGS__); \
    if (int st_ = ew_launch_check()) return st_;                                 \
  } while (0)

#define CHK(x) do { if (int st_ = (x)) return st_; } while (0)

// skewSymmetric / altSkewSymmetric (dns/nonlinear.cpp:36-411, Cartesian):
// on entry U holds the Fourier velocity; on exit U holds the physical-space
// velocity, Us[0] the Fourier copy and Uf[0] the Fourier N(u), smoothed.
int dns_nonlinear(semx_dns_s& D, hipStream_t s) {
  Mesh& M = *D.M;
  const size_t n = D.ntot;
  hipEvent_t ev;
  prof_begin(s, &ev);
  for (int i = 0; i < D.ncom; ++i) {
    if (D.nz > 2) {
      // the transform is out of place: its input stays as the Fourier copy
      // Us[0] and the previous Us[0] buffer becomes the spare (no copy)
      double* f = D.U[i];
      CHK(dft_launch(f, D.spare, D.nz, M.planeSize, -1, s));
      D.U[i] = D.spare;
      D.spare = D.Us[0][i];
      D.Us[0][i] = f;
    } else {
      EW(k_copy, D.Us[0][i], D.U[i]);
      CHK(dns_transform(D, D.U[i], -1, s));
    }
  }
  // skewSymmetric (0): both forms, halved; altSkewSymmetric (1): one form
  // per step in turn; convective (2): u . grad u; Stokes (3): none
  // (dns/nonlinear.cpp:36-211, 214-411, 414-525, 528-570)
  const bool conv = D.advection == 0 || D.advection == 2 || (D.advection == 1 && D.toggle);
  const bool cons = D.advection == 0 || (D.advection == 1 && !D.toggle);
  for (int i = 0; i < D.ncom; ++i) {
    double* N = D.Uf[0][i];
    bool zdefer = false;   // the conservative z term waits in tmp (Fourier space)
    double start = 0.0;    // the first x/y term stores N instead of updating it
    for (int j = 0; j < D.ndim; ++j) {
      if (conv) {   // n_i -= u_j d(u_i)/dx_j
        if (j == 2) {
          dim3 g((M.planeSize + 255) / 256, D.nz / 2);
          hipLaunchKernelGGL(k_grad_z, g, dim3(256), 0, s, D.Us[0][i], D.tmp, D.nz / 2,
                             M.planeSize, D.beta);
          CHK(ew_launch_check());
          CHK(dns_transform(D, D.tmp, -1, s));
          EW(k_times_minus, N, D.U[j], D.tmp);
        } else {   // fused: no copy of u_i, the product in the epilogue
          CHK(grad_nl_launch(M, D.nz, N, D.U[i], D.U[j], j, 0, s, 0.0, start));
          start = 1.0;
        }
      }
      if (cons) {   // n_i -= d(u_i u_j)/dx_j
        if (j == 2) {
          // d(u_i w)/dz is formed in Fourier space and, the transform being
          // linear, subtracted there after N's own forward transform below
          // (one inverse transform fewer than subtracting it in physical
          // space); tmp holds it until then
          EW(k_times, D.tmp, D.U[i], D.U[j]);
          CHK(dns_transform(D, D.tmp, +1, s));
          CHK(grad_launch(M, D.nz, D.beta, D.tmp, j, s));
          zdefer = true;
        } else {   // fused: the product formed on load, the difference in the epilogue
          CHK(grad_nl_launch(M, D.nz, N, D.U[i], D.U[j], j, 1, s, 0.0, start));
          start = 1.0;
        }
      }
    }
    if (!conv && !cons) EW(k_fill, N, 0.0);   // Stokes: no nonlinear terms
    const double half = D.advection == 0 ? 0.5 : 1.0;
    // the skew-symmetric halving, exact in binary and commuting with the
    // linear transform, rides on the deferred z term when there is one
    const bool fold = D.advection == 0 && zdefer;
    if (D.advection == 0 && !fold) EW(k_scale, N, 0.5);
    CHK(dns_transform(D, D.Uf[0][i], +1, s));
    if (fold) EW(k_axpby, D.Uf[0][i], 0.5, -half, D.tmp);
    else if (zdefer) EW(k_axpy, D.Uf[0][i], -half, D.tmp);
    CHK(smooth_launch(M, D.nz, D.Uf[0][i], s));
  }
  if (D.advection == 1) D.toggle = 1 - D.toggle;
  prof_end(s, "dns_nonlinear", ev);
  return SEMX_OK;
}

// BCmgr::maintainFourier + the pressure's evaluateBoundaries
// (dns/integrate.cpp:168-172) for step D.step (already advanced): Us[0]
// and Uf[0] hold the Fourier velocity and nonlinear terms of the step.
int dns_hopbc(semx_dns_s& D, hipStream_t s) {
  Mesh& M = *D.M;
  HopArgs A{};
  for (int i = 0; i < 3; ++i) A.U[i] = D.Us[0][i];
  A.N[0] = D.Uf[0][0]; A.N[1] = D.Uf[0][1];
  A.un = D.un; A.hop = D.hop; A.bcp = D.bcp;
  A.bel = D.bint; A.bsd = D.bint + M.nbound; A.pnat = D.bint + 2 * M.nbound;
  A.nrm = D.nrm;
  A.G = M.d.G; A.eset = M.d.eset; A.ops = M.d.ops;
  A.np = M.np; A.nz = D.nz; A.nline = D.nline; A.planeSize = M.planeSize;
  A.nord = D.nord; A.w3 = D.ncom == 3;
  A.Jd = std::min(D.step - 1, D.nord);
  A.Je = std::min(D.step, D.nord);
  if (A.Jd) stiffly_stable(A.Jd, A.a);
  extrapolation(A.Je, A.b);
  A.beta = D.beta; A.nu = D.kinvis; A.invdt = 1.0 / D.dt;
  const int nmodes = D.nz > 1 ? D.nz / 2 : 1;
  hipLaunchKernelGGL(k_hopbc, dim3(M.nbound, nmodes), dim3(256), 0, s, A);
  HIPCHK(hipGetLastError());
  return SEMX_OK;
}

// rollm (src/sem.h:106-120): level q takes level q-1, level 0 the last.
void roll(double* lev[3][3], int nord, int ncom) {
  if (nord < 2) return;
  for (int j = 0; j < ncom; ++j) {
    double* t = lev[nord - 1][j];
    for (int i = nord - 1; i; --i) lev[i][j] = lev[i - 1][j];
    lev[0][j] = t;
  }
}

// One time step of integrate (dns/integrate.cpp:79-221).
int dns_step(semx_dns_s& D, hipStream_t s) {
  Mesh& M = *D.M;
  const size_t n = D.ntot;
  const int ncom = D.ncom, nord = D.nord;
  CHK(dns_nonlinear(D, s));
  D.step += 1;
  if (D.Mp) CHK(dns_hopbc(D, s));
  const int Je = std::min(D.step, nord);
  double alpha[4], beta[3];
  stiffly_stable(Je, alpha);
  extrapolation(Je, beta);
  for (int q = 0; q < Je; ++q) beta[q] *= D.dt;
  // waveProp
  for (int i = 0; i < ncom; ++i) {
    WaveArgs w{};
    w.nlev = Je;
    for (int q = 0; q < Je; ++q) {
      w.us[q] = D.Us[q][i];
      w.uf[q] = D.Uf[q][i];
      w.a[q] = -alpha[q + 1];
      w.b[q] = beta[q];
    }
    EW(k_wave, D.U[i], w);
  }
  // swapData (U[i], Us[0][i]) and roll of the forcing levels
  for (int i = 0; i < ncom; ++i) std::swap(D.U[i], D.Us[0][i]);
  roll(D.Uf, nord, ncom);
  // setPForce: Uf[0][i] = d(Us[0][i])/dx_i, read out of place
  for (int i = 0; i < D.ndim; ++i) {
    if (i < 2) {
      CHK(grad_nl_launch(M, D.nz, D.Uf[0][i], D.Us[0][i], nullptr, i, 2, s));
    } else {
      dim3 g((M.planeSize + 255) / 256, D.nz / 2);
      hipLaunchKernelGGL(k_grad_z, g, dim3(256), 0, s, D.Us[0][i], D.Uf[0][i], D.nz / 2,
                         M.planeSize, D.beta);
      CHK(ew_launch_check());
    }
  }
  EW(k_divergence_sum, D.Uf[0][0], D.Uf[0][1], D.ndim > 2 ? D.Uf[0][2] : nullptr, D.dt);
  // pressure solve
  CHK(helm_solve(*D.pre, D.Uf[0][0], D.bcp, D.P, D.tol_rel, D.step_max, nullptr, s));
  // project
  {
    const double a = -1.0 / (D.dt * D.kinvis), b = 1.0 / D.kinvis;
    for (int i = 0; i < ncom; ++i) {
      std::swap(D.Us[0][i], D.Uf[0][i]);
      if (i >= D.ndim) EW(k_scale, D.Uf[0][i], a);
    }
    for (int i = 0; i < D.ndim; ++i) {   // Uf[0][i] = a Uf[0][i] + b dP/dx_i
      if (i < 2) {
        CHK(grad_nl_launch(M, D.nz, D.Uf[0][i], D.P, nullptr, i, 3, s, b, a));
      } else {
        dim3 g((M.planeSize + 255) / 256, D.nz / 2);
        hipLaunchKernelGGL(k_grad_z_axpby, g, dim3(256), 0, s, D.P, D.Uf[0][i], D.nz / 2,
                           M.planeSize, D.beta, a, b);
        CHK(ew_launch_check());
      }
    }
  }
  for (int i = 0; i < ncom; ++i) EW(k_copy, D.Us[0][i], D.U[i]);
  roll(D.Us, nord, ncom);
  // viscous solves: start-up JACPCG systems while step < N_TIME
  // (the velocity components share one system: one sweep of its factor)
  Helm* V = (D.step < nord) ? D.start[D.step - 1] : D.vel;
  double* fs[3] = {D.Uf[0][0], D.Uf[0][1], D.Uf[0][2]};
  double* us[3] = {D.U[0], D.U[1], D.U[2]};
  const double* bcs[3] = {D.bcu[0], D.bcu[1], D.bcu[2]};
  CHK(helm_solve_fields(*V, ncom, fs, D.Mp ? bcs : nullptr, us, D.tol_rel, D.step_max,
                        nullptr, s));
  return SEMX_OK;
}

DnsRoles roles_of(const semx_dns_s& D) {
  DnsRoles r;
  std::memset(&r, 0, sizeof(r));
  for (int i = 0; i < 3; ++i) r.U[i] = D.U[i];
  r.P = D.P;
  for (int q = 0; q < 3; ++q)
    for (int i = 0; i < 3; ++i) { r.Us[q][i] = D.Us[q][i]; r.Uf[q][i] = D.Uf[q][i]; }
  r.tmp = D.tmp;
  r.spare = D.spare;
  r.toggle = D.toggle;
  return r;
}

void set_roles(semx_dns_s& D, const DnsRoles& r) {
  for (int i = 0; i < 3; ++i) D.U[i] = r.U[i];
  D.P = r.P;
  for (int q = 0; q < 3; ++q)
    for (int i = 0; i < 3; ++i) { D.Us[q][i] = r.Us[q][i]; D.Uf[q][i] = r.Uf[q][i]; }
  D.tmp = r.tmp;
  D.spare = r.spare;
  D.toggle = r.toggle;
}

// A step from a HIP graph.  Start-up steps (order < N_TIME, JACPCG
// systems) and the first steady step (which runs the band-sweep autotuner
// and allocates the solve scratch) run eagerly, as do all steps while
// kernel timing is on.  A steady step whose role assignment has not been
// met is captured on the library's capture stream (the host state
// advances as in an eager step), instantiated and launched on the caller's
// stream; later steps with that assignment replay it.  Every kernel of the
// step takes its buffers by value, so one graph per assignment is exact.
// If a capture fails, graphs are switched off and the step runs eagerly.
int dns_step_graph(semx_dns_s& D, hipStream_t s) {
  if (!D.graphs_on || D.step < D.nord || prof_enabled()) return dns_step(D, s);
  const DnsRoles cur = roles_of(D);
  for (const DnsGraph& g : D.graphs) {
    if (g.before == cur) {
      HIPCHK(hipGraphLaunch(g.exec, s));
      set_roles(D, g.after);
      D.step += 1;
      D.replays += 1;
      return SEMX_OK;
    }
  }
  if (!D.capture) HIPCHK(hipStreamCreateWithFlags(&D.capture, hipStreamNonBlocking));
  const int step0 = D.step;
  HIPCHK(hipStreamBeginCapture(D.capture, hipStreamCaptureModeRelaxed));
  const int st = dns_step(D, D.capture);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(D.capture, &graph);
  hipGraphExec_t exec = nullptr;
  hipError_t ei = hipErrorUnknown;
  if (!st && ec == hipSuccess && graph) ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (st || ec != hipSuccess || ei != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "semx: DNS step capture failed (%s); stepping without graphs\n",
            st ? semx_error_string(st) : hipGetErrorString(ec != hipSuccess ? ec : ei));
    D.graphs_on = false;
    set_roles(D, cur);
    D.step = step0;
    return dns_step(D, s);
  }
  D.graphs.push_back({cur, roles_of(D), exec});
  HIPCHK(hipGraphLaunch(exec, s));
  D.replays += 1;
  return SEMX_OK;
}

void dns_free(semx_dns_s* D) {
  if (!D) return;
  for (DnsGraph& g : D->graphs) (void)hipGraphExecDestroy(g.exec);
  if (D->capture) (void)hipStreamDestroy(D->capture);
  for (Helm* h : {D->vel, D->pre, D->start[0], D->start[1]}) {
    if (!h) continue;
    helm_release(*h);
    delete h;
  }
  dev_free(D->store);
  dev_free(D->bstore);
  dev_free(D->bint);
  delete D;
}

int make_system(Mesh& M, double lambda2, double beta, int nz, int method, Helm** out) {
  Helm* H = new (std::nothrow) Helm();
  if (!H) return SEMX_E_NOMEM;
  H->M = &M;
  H->lambda2 = lambda2; H->beta = beta; H->base = 0;
  H->nmodes = nz > 1 ? nz / 2 : 1; H->nz = nz; H->method = method;
  int st = helm_build(*H);
  if (!st) st = helm_upload(*H);
  if (st) { helm_release(*H); delete H; return st; }
  *out = H;
  return SEMX_OK;
}

}  // namespace

// --- device memory of the allocation-free transform -------------------------
// semx_dft_r_noalloc works in this array: the transform matrix (at most
// 256 x 256 values, for tlen <= 256) and then the transformed field (at most
// SEMX_DFT_NOALLOC_MAX values).  Every device the library runs on holds its
// own copy until the process ends.  Host code never uses the array's name
// (a host-side stand-in, not a device address): it uses the device address
// obtained once per device by hipGetSymbolAddress.
constexpr size_t NOALLOC_W = 256 * 256;
static __device__ __align__(256) double scratch[NOALLOC_W + SEMX_DFT_NOALLOC_MAX];

namespace {

constexpr int NOALLOC_DEVICES = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[NOALLOC_DEVICES] = {};

}  // namespace

namespace semx {

// The current device's address of the array, looked up on its first use
// on that device (semx_device_init, or the first allocation-free call).
int noalloc_scratch(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_DEVICES) return SEMX_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      fprintf(stderr, "semx: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return SEMX_E_HIP;
    }
    g_noalloc_ptr[dev] = (double*)ptr;
  }
  if (out) *out = g_noalloc_ptr[dev];
  return SEMX_OK;
}

}  // namespace semx

namespace {

// Arguments of a transform as dDFTr checks them (femlib/fourier.c:47-62);
// 1 = nothing to do.
int dft_args(const double* data, int tlen, int ntrn, int sign) {
  if (!data || tlen < 1 || ntrn < 0 || (sign != 1 && sign != -1)) return SEMX_E_ARG;
  if (tlen < 2 || ntrn == 0) return -1;
  if (!fft_length_ok(tlen) || (ntrn & 1)) return SEMX_E_ARG;
  return SEMX_OK;
}

}  // namespace

extern "C" {

int semx_dft_r(double* data, int tlen, int ntrn, int sign, void* stream) {
  int st = dft_args(data, tlen, ntrn, sign);
  if (st) return st < 0 ? SEMX_OK : st;
  hipStream_t s = (hipStream_t)stream;
  static double* stage = nullptr;
  static size_t cap = 0;
  const size_t need = (size_t)tlen * ntrn;
  if (need > cap) {
    HIPCHK(hipStreamSynchronize(s));
    dev_free(stage);
    st = dev_upload<double>(&stage, nullptr, need);
    if (st) { cap = 0; return st; }
    cap = need;
  }
  st = dft_launch(data, stage, tlen, ntrn, sign, s);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(data, stage, need * sizeof(double), hipMemcpyDeviceToDevice, s));
  return SEMX_OK;
}

int semx_dft_r_noalloc(double* data, int tlen, int ntrn, int sign, void* stream) {
  int st = dft_args(data, tlen, ntrn, sign);
  if (st) return st < 0 ? SEMX_OK : st;
  const size_t need = (size_t)tlen * ntrn;
  if (tlen > 256 || need > SEMX_DFT_NOALLOC_MAX) {
    fprintf(stderr,
            "semx_dft_r_noalloc: %d x %d values exceed the static scratch "
            "(tlen <= 256, tlen*ntrn <= %zu)\n",
            tlen, ntrn, (size_t)SEMX_DFT_NOALLOC_MAX);
    return SEMX_E_SCRATCH;
  }
  double* base = nullptr;
  if ((st = noalloc_scratch(&base))) return st;
  hipStream_t s = (hipStream_t)stream;
  const int rows = 16 * dft_rp(tlen);
  const int ldw = (tlen + DFT_KB - 1) / DFT_KB * DFT_KB;
  double* W = base;
  double* out = base + NOALLOC_W;
  hipLaunchKernelGGL(k_dft_matrix, dim3((rows * ldw + 255) / 256), dim3(256), 0, s, W, tlen,
                     sign, rows, ldw);
  HIPCHK(hipGetLastError());
  if ((st = dft_launch_w(data, out, W, ldw, tlen, ntrn, s))) return st;
  HIPCHK(hipMemcpyAsync(data, out, need * sizeof(double), hipMemcpyDeviceToDevice, s));
  return SEMX_OK;
}

int semx_gradient(semx_mesh mesh, int nz, double beta, double* data, int dir, void* stream) {
  if (!mesh || !data || nz < 1 || dir < 0 || dir > 2) return SEMX_E_ARG;
  if (dir == 2 && (nz < 2 || (nz & 1))) return SEMX_E_ARG;
  Mesh& M = mesh->M;
  int st = mesh_upload(M);
  if (st) return st;
  return grad_launch(M, nz, beta, data, dir, (hipStream_t)stream);
}

int semx_couple(double* v, double* w, int nz, int n, int base_mode, int dir, void* stream) {
  if (!v || !w || nz < 1 || n < 1 || base_mode < 0 || (dir != 1 && dir != -1)) return SEMX_E_ARG;
  if (nz > 1 && (nz & 1)) return SEMX_E_ARG;
  const int nmode = nz / 2, klo = base_mode == 0 ? 1 : 0;   // 2D (nz = 1): nothing
  if (nmode <= klo) return SEMX_OK;
  const dim3 grid((unsigned)(((size_t)n + 255) / 256), (unsigned)(nmode - klo));
  hipLaunchKernelGGL(k_couple, grid, dim3(256), 0, (hipStream_t)stream, v, w, (size_t)n, klo,
                     nmode, dir);
  HIPCHK(hipGetLastError());
  return SEMX_OK;
}

int semx_smooth(semx_mesh mesh, int nz, double* data, void* stream) {
  if (!mesh || !data || nz < 1) return SEMX_E_ARG;
  Mesh& M = mesh->M;
  int st = mesh_upload(M);
  if (st) return st;
  return smooth_launch(M, nz, data, (hipStream_t)stream);
}

}  // extern "C"

namespace {

// The outward unit normals of every boundary side at its nodes, CCW
// (Element::sideGeom, src/element.cpp:1653-1744): side 0 -grad s, side 1
// grad r, side 2 grad s, side 3 -grad r, normalised.
std::vector<double> side_normals(const Mesh& M) {
  const int np = M.np, npnp = M.npnp, nline = M.nbound * np;
  std::vector<double> nrm(2 * (size_t)nline);
  for (int b = 0; b < M.nbound; ++b) {
    const int e = M.belmt[b], sd = M.bside[b];
    const double* g = M.G.data() + (size_t)M.eset[e] * G_N * npnp;
    for (int q = 0; q < np; ++q) {
      const int n = sd == 0 ? q : sd == 1 ? q * np + np - 1
                  : sd == 2 ? (np - 1) * np + (np - 1 - q) : (np - 1 - q) * np;
      const int rs = (sd == 0 || sd == 2) ? 1 : 0;     // s or r gradient
      const double sg = (sd == 0 || sd == 3) ? -1.0 : 1.0;
      const double a = sg * g[(rs ? G_DSDX : G_DRDX) * npnp + n];
      const double c = sg * g[(rs ? G_DSDY : G_DRDY) * npnp + n];
      const double len = std::hypot(a, c);
      nrm[b * np + q] = a / len;
      nrm[nline + b * np + q] = c / len;
    }
  }
  return nrm;
}

// Device storage of a bounded DNS: velocity BC lines copied from the
// caller's (device) lines, zeroed pressure lines and flux history, normals
// and the boundary tables.
int dns_bounded(semx_dns_s& D, Mesh& Mp, const double* const* bc) {
  Mesh& M = *D.M;
  D.Mp = &Mp;
  D.nline = M.nbound * M.np;
  const size_t nl = (size_t)D.nz * D.nline;
  const size_t total = (3 + 1 + 2 * (size_t)D.nord) * nl + 2 * (size_t)D.nline;
  CHK(dev_upload<double>(&D.bstore, nullptr, total));
  double* f = D.bstore;
  for (int i = 0; i < 3; ++i) { D.bcu[i] = f; f += nl; }
  D.bcp = f; f += nl;
  D.un = f; f += D.nord * nl;
  D.hop = f; f += D.nord * nl;
  D.nrm = f;
  for (int i = 0; i < D.ncom; ++i)
    HIPCHK(hipMemcpy(D.bcu[i], bc[i], nl * sizeof(double), hipMemcpyDeviceToDevice));
  const std::vector<double> nrm = side_normals(M);
  HIPCHK(hipMemcpy(D.nrm, nrm.data(), nrm.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<int> bint(3 * (size_t)M.nbound);
  for (int b = 0; b < M.nbound; ++b) {
    bint[b] = M.belmt[b];
    bint[M.nbound + b] = M.bside[b];
    bint[2 * M.nbound + b] = Mp.bkind[b] == 2;
  }
  return dev_upload<int>(&D.bint, bint.data(), bint.size());
}

int dns_create(Mesh& M, Mesh* Mp, const double* const* bc, int nz, double beta, double dt,
               double kinvis, int norder, int advection, double tol_rel, int step_max,
               semx_dns* dns) {
  if (!dns || nz < 1 || (nz > 1 && (nz & 1)) || norder < 1 || norder > 3 ||
      advection < 0 || advection > 3 || !(dt > 0.0) || !(kinvis > 0.0) || step_max < 1)
    return SEMX_E_ARG;
  if (nz > 2 && !fft_length_ok(nz)) return SEMX_E_ARG;
  *dns = nullptr;
  semx_dns_s* D = new (std::nothrow) semx_dns_s();
  if (!D) return SEMX_E_NOMEM;
  D->M = &M; D->nz = nz; D->nord = norder; D->advection = advection; D->beta = beta;
  D->dt = dt; D->kinvis = kinvis; D->tol_rel = tol_rel; D->step_max = step_max;
  D->ncom = nz > 1 ? 3 : 2; D->ndim = D->ncom;
  D->ntot = (size_t)nz * M.planeSize;
  int st = mesh_upload(M);
  const size_t nfield = 3 + 1 + 2 * 3 * 3 + 2;
  if (!st) st = dev_upload<double>(&D->store, nullptr, nfield * D->ntot);
  if (st) { dns_free(D); return st; }
  double* f = D->store;
  for (int i = 0; i < 3; ++i) D->U[i] = f + (size_t)i * D->ntot;
  f += 3 * D->ntot;
  D->P = f; f += D->ntot;
  for (int q = 0; q < 3; ++q)
    for (int i = 0; i < 3; ++i) {
      D->Us[q][i] = f; f += D->ntot;
      D->Uf[q][i] = f; f += D->ntot;
    }
  D->tmp = f; f += D->ntot;
  D->spare = f;
  double alpha[4];
  stiffly_stable(norder, alpha);
  st = make_system(M, alpha[0] / (dt * kinvis), beta, nz, SEMX_DIRECT, &D->vel);
  if (!st && Mp) st = mesh_upload(*Mp);
  if (!st && Mp) st = dns_bounded(*D, *Mp, bc);
  if (!st) st = make_system(Mp ? *Mp : M, 0.0, beta, nz, SEMX_DIRECT, &D->pre);
  for (int q = 1; q < norder && !st; ++q) {
    stiffly_stable(q, alpha);
    st = make_system(M, alpha[0] / (dt * kinvis), beta, nz, SEMX_JACPCG, &D->start[q - 1]);
  }
  if (st) { dns_free(D); return st; }
  *dns = D;
  return SEMX_OK;
}

}  // namespace

extern "C" {

int semx_dns_create(semx_mesh mesh, int nz, double beta, double dt, double kinvis, int norder,
                    int advection, double tol_rel, int step_max, semx_dns* dns) {
  if (!mesh) return SEMX_E_ARG;
  // a bounded mesh needs its BC lines and pressure numbering:
  // semx_dns_create_bc; cylindrical coordinates are not supported by the
  // time step
  if (mesh->M.nbound || mesh->M.cyl) return SEMX_E_UNSUPPORTED;
  return dns_create(mesh->M, nullptr, nullptr, nz, beta, dt, kinvis, norder, advection, tol_rel,
                    step_max, dns);
}

int semx_dns_create_bc(semx_mesh vmesh, semx_mesh pmesh, int nz, double beta, double dt,
                       double kinvis, int norder, int advection, double tol_rel, int step_max,
                       const double* bcu, const double* bcv, const double* bcw, semx_dns* dns) {
  if (!vmesh || !pmesh || !bcu || !bcv || (nz > 1 && !bcw)) return SEMX_E_ARG;
  const Mesh& M = vmesh->M;
  const Mesh& P = pmesh->M;
  if (M.cyl || P.cyl) return SEMX_E_UNSUPPORTED;
  // the two numberings must be of one element mesh with one boundary list
  if (!M.nbound || P.np != M.np || P.nel != M.nel || P.nbound != M.nbound ||
      P.belmt != M.belmt || P.bside != M.bside || P.x != M.x || P.y != M.y)
    return SEMX_E_MESH;
  const double* bc[3] = {bcu, bcv, bcw};
  return dns_create(vmesh->M, &pmesh->M, bc, nz, beta, dt, kinvis, norder, advection, tol_rel,
                    step_max, dns);
}

int semx_dns_set_iterative(semx_dns dns, int level) {
  if (!dns || dns->step || level < 0 || level > 2) return SEMX_E_ARG;
  semx_dns_s& D = *dns;
  double alpha[4];
  stiffly_stable(D.nord, alpha);
  Helm* vel = nullptr;
  Helm* pre = nullptr;
  int st = make_system(*D.M, alpha[0] / (D.dt * D.kinvis), D.beta, D.nz,
                       level >= 1 ? SEMX_JACPCG : SEMX_DIRECT, &vel);
  if (!st) st = make_system(D.Mp ? *D.Mp : *D.M, 0.0, D.beta, D.nz,
                            level >= 2 ? SEMX_MIXED : SEMX_DIRECT, &pre);
  if (st) {
    for (Helm* h : {vel, pre})
      if (h) { helm_release(*h); delete h; }
    return st;
  }
  for (Helm* h : {D.vel, D.pre}) { helm_release(*h); delete h; }
  D.vel = vel;
  D.pre = pre;
  return SEMX_OK;
}

int semx_dns_set_bc(semx_dns dns, const double* bcu, const double* bcv, const double* bcw,
                    void* stream) {
  if (!dns || !dns->Mp || !bcu || !bcv || (dns->ncom == 3 && !bcw)) return SEMX_E_ARG;
  const double* src[3] = {bcu, bcv, bcw};
  const size_t nl = (size_t)dns->nz * dns->nline;
  for (int i = 0; i < dns->ncom; ++i)
    HIPCHK(hipMemcpyAsync(dns->bcu[i], src[i], nl * sizeof(double), hipMemcpyDeviceToDevice,
                          (hipStream_t)stream));
  return SEMX_OK;
}

int semx_dns_destroy(semx_dns dns) {
  if (!dns) return SEMX_E_ARG;
  dns_free(dns);
  return SEMX_OK;
}

int semx_dns_run(semx_dns dns, int nstep, double* u, double* v, double* w, double* p,
                 void* stream) {
  if (!dns || nstep < 0 || !u || !v || !p || (dns->ncom == 3 && !w)) return SEMX_E_ARG;
  semx_dns_s& D = *dns;
  hipStream_t s = (hipStream_t)stream;
  double* ext[3] = {u, v, w};
  const size_t bytes = D.ntot * sizeof(double);
  for (int i = 0; i < D.ncom; ++i)
    HIPCHK(hipMemcpyAsync(D.U[i], ext[i], bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(D.P, p, bytes, hipMemcpyDeviceToDevice, s));
  for (int k = 0; k < nstep; ++k) CHK(dns_step_graph(D, s));
  for (int i = 0; i < D.ncom; ++i)
    HIPCHK(hipMemcpyAsync(ext[i], D.U[i], bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(p, D.P, bytes, hipMemcpyDeviceToDevice, s));
  // the solves of the steps ran without host synchronisation: any expired
  // cooperative hand-off wait among them is reported here
  int st = SEMX_OK;
  for (Helm* h : {D.vel, D.pre})
    if (h && !st) st = helm_check(*h, s);
  return st;
}

int semx_dns_info(semx_dns dns, int* info) {
  if (!dns || !info) return SEMX_E_ARG;
  info[0] = dns->step; info[1] = dns->nz; info[2] = dns->ncom; info[3] = dns->nord;
  return SEMX_OK;
}

int semx_dns_set_graph(semx_dns dns, int on) {
  if (!dns) return SEMX_E_ARG;
  dns->graphs_on = on != 0;
  return SEMX_OK;
}

int semx_dns_graph_info(semx_dns dns, int* ngraphs, long long* replays) {
  if (!dns || !ngraphs || !replays) return SEMX_E_ARG;
  *ngraphs = (int)dns->graphs.size();
  *replays = dns->replays;
  return SEMX_OK;
}

}  // extern "C"
