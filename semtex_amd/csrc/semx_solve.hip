// semx_solve.hip -- CDNA4 (gfx950) kernels and the C ABI of libsemx for
// semtex's Field::solve (src/field.cpp:509-725) over every plane of a field:
//
//   k_ess_gather    essential BC values of every plane   (Field::getEssential)
//   k_condense[_staged]
//                   -(M f + H g), interior elimination F_b -= hbi F_i, <h,w>
//                   (Field::constrain, Element::local2globalSumSC, Natural::sum)
//   k_assemble      owner-gather of the condensed boundary RHS (buildRHS)
//   k_band_solve    banded Cholesky triangular solves U^T y = b, U x = y
//                   (LAPACK dpbtrs): one workgroup per mode, the factor
//                   streamed as NB-column panels, the RHS in an LDS window
//   k_band_coop     the same with several workgroups per mode in a lagged
//                   pipeline (opt-in)
//   k_backsub[_staged]
//                   interior recovery u_i = hii F_i - hbi^T u_b and essential
//                   re-imposition (Element::global2localSC, Field::setEssential)
//   k_pcg           JACPCG, persistent workgroups, one or several per plane
//                   (FP64, or FP32 / BF16-matrix-core inner PCG with FP64
//                   refinement)
//   k_pcg_batch     the same for many independent 2D problems in one launch,
//                   right-hand side built in the problem's workgroups
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <type_traits>
#include <chrono>
#include <map>
#include <string>

#include "semx_hip.h"

using namespace semx;

namespace {

constexpr int ELT_THREADS = 1024;  // element kernels: several elements per block

struct ElemArgs {
  int np, npnp, next, nint, nel, planeSize, nline, nsolve, nzero, nset;
  double lambda2, beta;
  int base;
  const double* hcon;     // per mode: lambda2 + betak2 (1 with mq)
  const double* mq;       // cylindrical: per mode and set, the operator's mass
                          // factors (nmodes*nset*npnp); null: Q4 of the set
  const double* G;
  const int* eset;
  const int* emap;
  const int* pmap;
  const int* btog;
  const int* emask;
  const double* ops;      // DV DT SDV SDT
  const int* nat_ptr;
  const int* nat_b;
  const int* bside;
  const double* barea;
  const double* hbi;
  const double* hii;
  const int* plane_of;
  const int* mode_n;
  const int* mode_kind;   // per mode: SEMX_DIRECT or SEMX_JACPCG
  double* Fb;
  double* Fi;
  double* g;
  const double* sol;
  int solstride;
  const double* f;
  const double* bc;
  double* u;
};

__global__ void k_ess_gather(const double* __restrict__ bc, const int* __restrict__ ess_src,
                             const int* __restrict__ plane_of, double* __restrict__ g,
                             int nzero, int nline) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (m >= nzero) return;
  const int k = plane_of[p];
  g[(size_t)p * nzero + m] = bc[(size_t)k * nline + ess_src[m]];
}

// Elements of one plane per block: CE = blockDim / (np*np) elements, thread
// t works on element t / npnp, node t % npnp.
// One chunk of ce elements [e0, min(e0 + ce, e_end)) of plane-list entry p.
// sHbiT: the mode's hbi staged transposed in LDS ([c][b]), or null to read
// hbi per element from global memory.  Called uniformly by the block.
__device__ inline void condense_chunk(const ElemArgs& a, int p, int e0, int e_end,
                                      const double* sHbiT, double* esh) {
  const int t = threadIdx.x;
  const int k = a.plane_of[p];
  const int np = a.np, npnp = a.npnp, next = a.next, nint = a.nint;
  const int ce = ELT_THREADS / npnp;
  const int el = t / npnp, pt = t - el * npnp;
  const int e = e0 + el;
  const bool act = el < ce && e < e_end;
  const int m = k >> 1;
  const double hcon = a.hcon[m];
  const bool svv = a.lambda2 > EPSDP;
  double* sD = esh;                       // npnp
  double* sT = sD + npnp;                 // npnp
  double* sU = sT + npnp;                 // ce*npnp
  double* sR = sU + ce * npnp;
  double* sS = sR + ce * npnp;
  double* sF = sS + ce * npnp;
  __syncthreads();   // the previous chunk's reads of the buffers are done
  if (t < npnp) {
    sD[t] = a.ops[(svv ? 2 : 0) * npnp + t];
    sT[t] = a.ops[(svv ? 3 : 1) * npnp + t];
  }
  const int set = act ? a.eset[e] : 0;
  const double* g = a.G + (size_t)set * G_N * npnp;
  double F = 0.0;
  if (act) F = g[G_Q4 * npnp + pt] * a.f[(size_t)k * a.planeSize + (size_t)e * npnp + pt];
  const int em = act ? a.emask[e] : 0;
  // -(M f + H g): the lifting H g only on elements with essential nodes
  if (__syncthreads_or(em)) {
    const int base = el * npnp;
    if (act) sU[base + pt] = 0.0;
    __syncthreads();
    if (act && em && pt < next) {
      const int gid = a.btog[e * next + pt];
      sU[base + a.emap[pt]] =
          (gid >= a.nsolve) ? a.g[(size_t)p * a.nzero + (gid - a.nsolve)] : 0.0;
    }
    __syncthreads();
    const int i = pt / np, j = pt - (pt / np) * np;
    if (act) {
      double r = 0.0, s2 = 0.0;
      for (int q = 0; q < np; ++q) {
        r += sU[base + i * np + q] * sT[q * np + j];
        s2 += sD[i * np + q] * sU[base + q * np + j];
      }
      const double q1 = g[G_Q1 * npnp + pt], q2 = g[G_Q2 * npnp + pt], q3 = g[G_Q3 * npnp + pt];
      sR[base + pt] = q1 * r + q3 * s2;
      sS[base + pt] = q2 * s2 + q3 * r;
    }
    __syncthreads();
    if (act && em) {
      const double* q4 = a.mq ? a.mq + ((size_t)m * a.nset + set) * npnp : g + G_Q4 * npnp;
      double T = q4[pt] * sU[base + pt] * hcon;
      double s1 = 0.0, s3 = 0.0;
      for (int q = 0; q < np; ++q) {
        s1 += sT[i * np + q] * sS[base + q * np + j];
        s3 += sR[base + i * np + q] * sD[q * np + j];
      }
      T += s1;
      T += s3;
      F += T;
    }
  }
  // element values in emap order: boundary first, then interior
  if (act) sF[el * npnp + a.pmap[pt]] = -F;
  __syncthreads();
  // interior part out; boundary part condensed: F_b - hbi F_i, plus <h,w>
  for (int idx = t; idx < ce * npnp; idx += ELT_THREADS) {
    const int le = idx / npnp, q = idx - le * npnp;
    const int ee = e0 + le;
    if (ee >= e_end) continue;
    const size_t pe = (size_t)p * a.nel + ee;
    const double* Fe = sF + le * npnp;
    if (q >= next) {
      a.Fi[pe * nint + (q - next)] = Fe[q];
    } else {
      const int b = q;
      const double* Fi = Fe + next;
      double s0 = 0.0, s1 = 0.0;
      if (a.mode_kind[m] == SEMX_DIRECT) {  // PCG modes keep the raw F_b
        if (sHbiT) {
          int c = 0;
          for (; c + 1 < nint; c += 2) {
            s0 += sHbiT[c * next + b] * Fi[c];
            s1 += sHbiT[(c + 1) * next + b] * Fi[c + 1];
          }
          if (c < nint) s0 += sHbiT[c * next + b] * Fi[c];
        } else {
          const double* hbi = a.hbi + ((size_t)m * a.nset + a.eset[ee]) * next * nint;
          int c = 0;
          for (; c + 1 < nint; c += 2) {
            s0 += hbi[b * nint + c] * Fi[c];
            s1 += hbi[b * nint + c + 1] * Fi[c + 1];
          }
          if (c < nint) s0 += hbi[b * nint + c] * Fi[c];
        }
      }
      double fb = Fe[b] - (s0 + s1);
      const int nm = np - 1;
      for (int cc = a.nat_ptr[ee]; cc < a.nat_ptr[ee + 1]; ++cc) {
        const int bb = a.nat_b[cc];
        const int side = a.bside[bb];
        const int ii = (b - side * nm + next) % next;
        if (ii <= nm) fb += a.bc[(size_t)k * a.nline + bb * np + ii] * a.barea[bb * np + ii];
      }
      a.Fb[pe * next + b] = fb;
    }
  }
}

// Elements of one plane per block: ce = blockDim / (np*np) elements, thread
// t works on element t / npnp, node t % npnp.
__global__ __launch_bounds__(ELT_THREADS) void k_condense(ElemArgs a) {
  extern __shared__ double esh[];
  const int ce = ELT_THREADS / a.npnp;
  condense_chunk(a, blockIdx.y, blockIdx.x * ce, a.nel, nullptr, esh);
}

// Synchronisation of one wavefront's LDS traffic.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Meshes with one geometric set: one block per (mode, group of ge
// elements), the mode's hbi staged transposed in LDS.  Each wavefront forms
// the forcing of one (element, plane) at a time in emap order -- mass term,
// and on elements with essential nodes the lifting -H g by the tensor-
// product operator in a private LDS slice with wavefront synchronisation
// only -- and after one block barrier all threads condense the block's
// elements, both planes: interior part out, F_b - hbi F_i plus the natural
// boundary terms.
constexpr int CND_WAVES = ELT_THREADS / 64;

__global__ __launch_bounds__(ELT_THREADS, 8) void k_condense_staged(ElemArgs a,
                                                                 const int* __restrict__ mode_plane,
                                                                 int ge) {
  extern __shared__ double esh[];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int np = a.np, next = a.next, nint = a.nint, npnp = a.npnp;
  const int m = blockIdx.y;
  const bool direct = a.mode_kind[m] == SEMX_DIRECT;
  const double hcon = a.hcon[m];
  const bool svv = a.lambda2 > EPSDP;
  double* sHbiT = esh;                          // nint*next, [c][b]
  double* sD = sHbiT + (size_t)next * nint;     // npnp
  double* sT = sD + npnp;                       // npnp
  double* sF = sT + npnp;                       // 2*ge*npnp, [plane][element], emap order
  double* sW = sF + 2 * (size_t)ge * npnp;      // per wave 3*npnp: U, R, S
  if (direct) {
    const double* hbi = a.hbi + (size_t)m * next * nint;
    for (int i = t; i < next * nint; i += ELT_THREADS) {
      const int b = i / nint, c = i - (i / nint) * nint;
      sHbiT[c * next + b] = hbi[i];
    }
  }
  for (int i = t; i < npnp; i += ELT_THREADS) {
    sD[i] = a.ops[(svv ? 2 : 0) * npnp + i];
    sT[i] = a.ops[(svv ? 3 : 1) * npnp + i];
  }
  __syncthreads();
  const int e_begin = blockIdx.x * ge, e_end = min(a.nel, e_begin + ge);
  const int ne = e_end - e_begin;
  const int p0 = mode_plane[2 * m], p1 = mode_plane[2 * m + 1];
  const int npl = p1 >= 0 ? 2 : 1;
  const double* g = a.G;                        // the one geometric set
  double* sU = sW + (size_t)wave * 3 * npnp;
  double* sR = sU + npnp;
  double* sS = sR + npnp;
  for (int item = wave; item < ne * npl; item += CND_WAVES) {
    const int pl = item / ne, le = item - pl * ne;
    const int ee = e_begin + le;
    const int p = pl ? p1 : p0;
    const int k = a.plane_of[p];
    double* Fe = sF + (size_t)(pl * ge + le) * npnp;
    const bool em = a.emask[ee] != 0;           // uniform over the wavefront
    if (em) {
      for (int pt = lane; pt < npnp; pt += 64) sU[pt] = 0.0;
      wave_sync();
      for (int pt = lane; pt < next; pt += 64) {
        const int gid = a.btog[ee * next + pt];
        sU[a.emap[pt]] = (gid >= a.nsolve) ? a.g[(size_t)p * a.nzero + (gid - a.nsolve)] : 0.0;
      }
      wave_sync();
      for (int pt = lane; pt < npnp; pt += 64) {
        const int i = pt / np, j = pt - (pt / np) * np;
        double r = 0.0, s2 = 0.0;
        for (int q = 0; q < np; ++q) {
          r += sU[i * np + q] * sT[q * np + j];
          s2 += sD[i * np + q] * sU[q * np + j];
        }
        const double q1 = g[G_Q1 * npnp + pt], q2 = g[G_Q2 * npnp + pt], q3 = g[G_Q3 * npnp + pt];
        sR[pt] = q1 * r + q3 * s2;
        sS[pt] = q2 * s2 + q3 * r;
      }
      wave_sync();
    }
    for (int pt = lane; pt < npnp; pt += 64) {
      double F = g[G_Q4 * npnp + pt] * a.f[(size_t)k * a.planeSize + (size_t)ee * npnp + pt];
      if (em) {
        const int i = pt / np, j = pt - (pt / np) * np;
        const double* q4 = a.mq ? a.mq + (size_t)m * npnp : g + G_Q4 * npnp;
        double T = q4[pt] * sU[pt] * hcon;
        double s1 = 0.0, s3 = 0.0;
        for (int q = 0; q < np; ++q) {
          s1 += sT[i * np + q] * sS[q * np + j];
          s3 += sR[i * np + q] * sD[q * np + j];
        }
        T += s1;
        T += s3;
        F += T;
      }
      Fe[a.pmap[pt]] = -F;
    }
    wave_sync();   // this wavefront's slice is free for its next item
  }
  __syncthreads();
  const int nm = np - 1;
  for (int idx = t; idx < npl * ne * npnp; idx += ELT_THREADS) {
    const int it = idx / npnp, q = idx - it * npnp;
    const int pl = it / ne, le = it - pl * ne;
    const int ee = e_begin + le;
    const int p = pl ? p1 : p0;
    const int k = a.plane_of[p];
    const size_t pe = (size_t)p * a.nel + ee;
    const double* Fe = sF + (size_t)(pl * ge + le) * npnp;
    if (q >= next) {
      a.Fi[pe * nint + (q - next)] = Fe[q];
    } else {
      const int b = q;
      const double* Fi = Fe + next;
      double s0 = 0.0, s1 = 0.0;
      if (direct) {   // PCG modes keep the raw F_b
        int c = 0;
        for (; c + 1 < nint; c += 2) {
          s0 += sHbiT[c * next + b] * Fi[c];
          s1 += sHbiT[(c + 1) * next + b] * Fi[c + 1];
        }
        if (c < nint) s0 += sHbiT[c * next + b] * Fi[c];
      }
      double fb = Fe[b] - (s0 + s1);
      for (int cc = a.nat_ptr[ee]; cc < a.nat_ptr[ee + 1]; ++cc) {
        const int bb = a.nat_b[cc];
        const int side = a.bside[bb];
        const int ii = (b - side * nm + next) % next;
        if (ii <= nm) fb += a.bc[(size_t)k * a.nline + bb * np + ii] * a.barea[bb * np + ii];
      }
      a.Fb[pe * next + b] = fb;
    }
  }
}

struct BandArgs {
  const double* Pf;         // forward panels
  const double* Pb;         // backward panels
  const double* Linv;
  const double* LinvT;
  const int* mode_n;
  const long long* offP;
  const long long* offL;
  const int* dmode;
  const int* mode_plane;
  const int* asm_ptr;
  const int* asm_idx;
  const double* Fb;
  double* sol;
  int solstride, kd, kdp, nel, next, ring;
  int npl;                  // right-hand sides per field and mode (1 or 2 planes)
  long long fstride;        // sol offset between fields
  const int* xrows;         // extended panels: live rows per block and direction
  int xblk;                 // block stride of xrows per mode
  // right-hand sides split over fsplit workgroups per mode (each NR of
  // them), the workgroups of one mode on one XCD so that their reads of the
  // same panels share its L2; fsplit = 1: one workgroup per mode
  int fsplit = 1, ndm = 0;
};

// mode and right-hand-side group of this workgroup (md < 0: padding block).
// Dispatch is round-robin over the 8 XCDs, so blocks b and b + 8 k share one.
__device__ inline void band_block(const BandArgs& a, int& md, int& fsel) {
  const int b = blockIdx.x;
  if (a.fsplit <= 1) { md = b; fsel = 0; return; }
  const int F = a.fsplit;
  md = (b / (8 * F)) * 8 + (b & 7);
  fsel = (b >> 3) % F;
  if (md >= a.ndm) md = -1;
}

template <int G>
__device__ inline double group_sum(double v) {
  // sum over aligned groups of G lanes (G a power of two <= 32)
#pragma unroll
  for (int off = G / 2; off >= 1; off /= 2) v += __shfl_xor(v, off);
  return v;
}

// Condensed right-hand side of every plane: owner gather of the element-
// boundary contributions (the direct-stiffness sum of
// Element::local2globalSumSC), zero beyond the solved unknowns.
__global__ void k_assemble(const double* __restrict__ Fb, const int* __restrict__ asm_ptr,
                           const int* __restrict__ asm_idx, const int* __restrict__ plane_of,
                           const int* __restrict__ mode_n, const int* __restrict__ mode_kind,
                           double* __restrict__ sol, int solstride, int nglobal, int nelnext) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= nglobal) return;
  if (mode_kind[plane_of[p] >> 1] != SEMX_DIRECT) return;
  const int n = mode_n[plane_of[p] >> 1];
  double v = 0.0;
  if (i < n) {
    const double* fb = Fb + (size_t)p * nelnext;
    for (int k = asm_ptr[i]; k < asm_ptr[i + 1]; ++k) v += fb[asm_idx[k]];
  }
  sol[(size_t)p * solstride + i] = v;
}

// One sweep of the block-triangular band solve for NR right-hand sides.
// Only a window of the right-hand side lives in LDS: a ring of W (power of
// two >= kd + 2 NB) slots, index i in slot i & (W-1).  Entries are admitted
// from global memory (the assembled RHS, then the forward result) as the
// sweep reaches them, and each block's solved values go to global memory
// and to a small LDS staging buffer.
//   forward : y_J <- Linv_J y_J;  y[j0+NB+r] -= sum_jj Pf_J[jj][r] y[j0+jj]
//   backward: x_J <- LinvT_J y_J; y[j0-1-r]  -= sum_jj Pb_J[jj][r] x[j0+jj]
// Thread r owns update row r (NPASS rows per thread); the panel slice,
// inverse slice and admitted entries of block s+D are loaded into
// registers while block s is processed; two barriers per block.
template <int NB, int NPASS, int BT, int NR, int D, bool FWD>
__device__ inline void band_sweep(const double* __restrict__ P, const double* __restrict__ L,
                                  double* const (&ring)[NR], double* const (&gy)[NR],
                                  double* __restrict__ ybuf, int n, int nblk, int kd, int kdp,
                                  int wm) {
  constexpr int QPT = (NB * NB / BT > 0) ? NB * NB / BT : 1;  // diag products per thread
  constexpr int G = NB / QPT;                                 // lanes per diag row
  const int tid = threadIdx.x;
  const int npad = nblk * NB;
  const bool dact = tid < NB * G;
  const int da = dact ? tid / G : 0, dg = dact ? tid % G : 0;
  double pbuf[D + 1][NPASS][NB], lbuf[D + 1][QPT], abuf[D + 1][NR];

  // first window: forward [0, min(npad, NB+kd)), backward [max(0,j0-kd), npad)
  {
    const int lo = FWD ? 0 : max(0, (nblk - 1) * NB - kd);
    const int hi = FWD ? min(npad, NB + kd) : npad;
    for (int i = lo + tid; i < hi; i += BT) {
#pragma unroll
      for (int c = 0; c < NR; ++c) ring[c][i & wm] = (i < n) ? gy[c][i] : 0.0;
    }
  }
  // entries admitted at step s >= 1: forward [j0+kd, j0+kd+NB) below npad,
  // backward [j0-kd, j0-kd+NB) from 0 up; thread t < NB owns one
  auto admit_index = [&](int s) {
    const int J = FWD ? s : nblk - 1 - s;
    const int i = FWD ? J * NB + kd + tid : J * NB - kd + tid;
    const bool ok = tid < NB && (FWD ? (i >= (NB + kd) && i < npad) : (i >= 0));
    return ok ? i : -1;
  };

  auto load = [&](int s, double (&pp)[NPASS][NB], double (&ll)[QPT], double (&ad)[NR]) {
    const int J = FWD ? s : nblk - 1 - s;
    const double* pj = P + (size_t)J * NB * kdp;
#pragma unroll
    for (int q = 0; q < NPASS; ++q) {
      const int r = min(tid + q * BT, kdp - 1);
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) pp[q][jj] = pj[(size_t)jj * kdp + r];
    }
    const double* lj = L + (size_t)J * NB * NB + da * NB + dg * QPT;
#pragma unroll
    for (int q = 0; q < QPT; ++q) ll[q] = lj[q];
    // unconditional load from a clamped index (the validity select is made
    // where the value is used, so no branch splits the load stream)
    const int ic = min(max(admit_index(s), 0), n - 1);
#pragma unroll
    for (int c = 0; c < NR; ++c) ad[c] = gy[c][ic];
  };

  auto step = [&](int s, double (&pp)[NPASS][NB], double (&ll)[QPT], double (&ad)[NR]) {
    const int J = FWD ? s : nblk - 1 - s;
    const int j0 = J * NB;
    const int r0 = j0 & wm;  // blocks never wrap: W is a multiple of NB
    const int ia = admit_index(s);
    if (s > 0 && ia >= 0) {
#pragma unroll
      for (int c = 0; c < NR; ++c) ring[c][ia & wm] = (ia < n) ? ad[c] : 0.0;
    }
    __syncthreads();
    double v[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < QPT; ++q) acc += ll[q] * ring[c][r0 + dg * QPT + q];
      v[c] = group_sum<G>(acc);
    }
    if (dact && dg == 0) {
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        ybuf[c * NB + da] = v[c];
        if (j0 + da < n) gy[c][j0 + da] = v[c];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NPASS; ++q) {
      const int r = tid + q * BT;
      const int i = FWD ? j0 + NB + r : j0 - 1 - r;
      if (r < kd && (FWD ? i < n : i >= 0)) {
        const int ri = i & wm;
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
          for (int jj = 0; jj < NB; jj += 4) {
            a0 += pp[q][jj] * ybuf[c * NB + jj];
            a1 += pp[q][jj + 1] * ybuf[c * NB + jj + 1];
            a2 += pp[q][jj + 2] * ybuf[c * NB + jj + 2];
            a3 += pp[q][jj + 3] * ybuf[c * NB + jj + 3];
          }
          ring[c][ri] -= (a0 + a1) + (a2 + a3);
        }
      }
    }
  };

  if (nblk <= 0) return;
  // Prefetches are unconditional (the step index is clamped, so the last
  // ones reload the final block): every path into a step then has the same
  // number of loads in flight and waits only for its own block's data.
#pragma unroll
  for (int d = 0; d < D; ++d) load(min(d, nblk - 1), pbuf[d], lbuf[d], abuf[d]);
  for (int s = 0; s < nblk; s += D + 1) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      if (s + u >= nblk) break;
      const int nx = (u + D) % (D + 1);
      load(min(s + u + D, nblk - 1), pbuf[nx], lbuf[nx], abuf[nx]);
      step(s + u, pbuf[u], lbuf[u], abuf[u]);
    }
  }
  // the backward sweep re-reads, from global memory, values that other
  // threads of this workgroup wrote during the forward sweep
  if (FWD) __threadfence();
  __syncthreads();
}

template <int NB, int NPASS, int BT, int NR, int D>
__global__ __launch_bounds__(BT) void k_band_solve(BandArgs a) {
  extern __shared__ double lds[];
  const int m = a.dmode[blockIdx.x];
  const int n = a.mode_n[m];
  const int nblk = (n + NB - 1) / NB;
  double* ring[NR];
  double* gy[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    // right-hand side c: plane c % npl of the mode in field c / npl; a mode
    // with fewer planes than npl repeats its first plane (the duplicate
    // sweeps identical data and writes identical values)
    const int fld = c / a.npl, kk = c - fld * a.npl;
    int p = a.mode_plane[2 * m + kk];
    if (p < 0) p = a.mode_plane[2 * m];
    ring[c] = lds + (size_t)c * a.ring;
    gy[c] = a.sol + fld * a.fstride + (size_t)p * a.solstride;
  }
  double* ybuf = lds + (size_t)NR * a.ring;
  band_sweep<NB, NPASS, BT, NR, D, true>(a.Pf + a.offP[m], a.Linv + a.offL[m], ring, gy,
                                          ybuf, n, nblk, a.kd, a.kdp, a.ring - 1);
  band_sweep<NB, NPASS, BT, NR, D, false>(a.Pb + a.offP[m], a.LinvT + a.offL[m], ring, gy,
                                           ybuf, n, nblk, a.kd, a.kdp, a.ring - 1);
}

// Band solve over extended panels, which carry the diagonal-block inverse
// folded in.  Panel of block J: NB columns of ldx rows, stored as column
// pairs (entry (kk, r) at ((kk/2) ldx + r) 2 + kk%2, one 16-byte load per
// row and pair, which keeps a step's loads within the wave's limit of
// outstanding memory instructions),
//   rows [0, NB)        Linv_J (forward) / LinvT_J (backward): row r gives
//                       y_J[r] (x_J[r]) from the window values of block J,
//   rows [NB, NB + kd)  Pf_J Linv_J (forward) / Pb_J LinvT_J (backward): the
//                       update of the kd rows after (before) the block.
// One step is a single product of the panel with the NB window values of
// block J: rows below NB are stored as the block's solution and the rest
// are subtracted from the window.  One barrier per step; the loads of a
// step are its panel slice and admitted entries only, prefetched D steps
// ahead into registers.
constexpr int XNB = 32;

template <int NB, int NPASS, int BT, int NR, int D, bool FWD>
__device__ inline void band_ext_sweep(const double* __restrict__ P,
                                      double* const (&ring)[NR], double* const (&gy)[NR],
                                      const int* __restrict__ xrows, int n, int nblk, int kd,
                                      int ldx, int wm) {
  constexpr int WC = (NR >= 4 || NB < 16) ? 4 : 8;   // window values per read batch
  const int tid = threadIdx.x;
  const int npad = nblk * NB;
  double pbuf[D + 1][NPASS][NB], abuf[D + 1][NR];
  {
    const int lo = FWD ? 0 : max(0, (nblk - 1) * NB - kd);
    const int hi = FWD ? min(npad, NB + kd) : npad;
    for (int i = lo + tid; i < hi; i += BT) {
#pragma unroll
      for (int c = 0; c < NR; ++c) ring[c][i & wm] = (i < n) ? gy[c][i] : 0.0;
    }
  }
  auto admit_index = [&](int s) {
    const int J = FWD ? s : nblk - 1 - s;
    const int i = FWD ? J * NB + kd + tid : J * NB - kd + tid;
    const bool ok = tid < NB && (FWD ? (i >= (NB + kd) && i < npad) : (i >= 0));
    return ok ? i : -1;
  };
  // rows past a block's live rows load the last live row's address (the
  // same cache lines, no extra memory traffic) and skip the product
  auto load = [&](int s, double (&pp)[NPASS][NB], double (&ad)[NR]) {
    const int J = FWD ? s : nblk - 1 - s;
    const double2* pj = reinterpret_cast<const double2*>(P + (size_t)J * NB * ldx);
    const int rlast = xrows[2 * J + (FWD ? 0 : 1)] - 1;
#pragma unroll
    for (int q = 0; q < NPASS; ++q) {
      const int r = min(tid + q * BT, rlast);
#pragma unroll
      for (int k2 = 0; k2 < NB / 2; ++k2) {
        const double2 v = pj[(size_t)k2 * ldx + r];
        pp[q][2 * k2] = v.x;
        pp[q][2 * k2 + 1] = v.y;
      }
    }
    const int ic = min(max(admit_index(s), 0), n - 1);
#pragma unroll
    for (int c = 0; c < NR; ++c) ad[c] = gy[c][ic];
  };
  auto step = [&](int s, double (&pp)[NPASS][NB], double (&ad)[NR]) {
    const int J = FWD ? s : nblk - 1 - s;
    const int j0 = J * NB;
    const int r0 = j0 & wm;  // blocks never wrap: W is a multiple of NB
    const int ia = admit_index(s);
    const int rows = xrows[2 * J + (FWD ? 0 : 1)];
    // Every thread takes delivery of this step's registers here, on every
    // path: consumed only inside the branches below, the loads would count
    // as possibly outstanding at the loop head, where the compiler would
    // then wait for all later prefetches as well.
#pragma unroll
    for (int q = 0; q < NPASS; ++q)
#pragma unroll
      for (int kk = 0; kk < NB; ++kk) asm volatile("" ::"v"(pp[q][kk]));
#pragma unroll
    for (int c = 0; c < NR; ++c) asm volatile("" ::"v"(ad[c]));
    if (s > 0 && ia >= 0) {
#pragma unroll
      for (int c = 0; c < NR; ++c) ring[c][ia & wm] = (ia < n) ? ad[c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NPASS; ++q) {
      const int r = tid + q * BT;
      const int rr = r - NB;
      const int i = r < NB ? j0 + r : (FWD ? j0 + NB + rr : j0 - 1 - rr);
      const bool live = r < NB ? i < n : (r < rows && rr < kd && (FWD ? i < n : i >= 0));
      if (live) {
        // one right-hand side at a time, its window read in batches of WC
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
          for (int h = 0; h < NB; h += WC) {
            const double2* wv = reinterpret_cast<const double2*>(ring[c] + r0 + h);
            double2 w[WC / 2];
#pragma unroll
            for (int k = 0; k < WC / 2; ++k) w[k] = wv[k];
#pragma unroll
            for (int k = 0; k < WC / 2; ++k) {
              acc0 += pp[q][h + 2 * k] * w[k].x;
              acc1 += pp[q][h + 2 * k + 1] * w[k].y;
            }
          }
          const double v = acc0 + acc1;
          if (r < NB) gy[c][i] = v;
          else ring[c][i & wm] -= v;
        }
      }
    }
  };

  if (nblk <= 0) return;
#pragma unroll
  for (int d = 0; d < D; ++d) load(min(d, nblk - 1), pbuf[d], abuf[d]);
  // whole groups of D+1 steps with no exit inside a group, then the rest
  int s = 0;
  for (; s + D < nblk; s += D + 1) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      const int nx = (u + D) % (D + 1);
      load(min(s + u + D, nblk - 1), pbuf[nx], abuf[nx]);
      step(s + u, pbuf[u], abuf[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < D; ++u)
    if (s + u < nblk) step(s + u, pbuf[u], abuf[u]);
  // the backward sweep re-reads, from global memory, values that other
  // threads of this workgroup wrote during the forward sweep
  if (FWD) __threadfence();
  __syncthreads();
}

// The same product over extended panels with the whole right-hand side
// resident in LDS (NR vectors of nblk*NB values), solved in place: the
// loops then touch global memory only for panel loads.  Two threads share a
// panel row (lane pair 2t, 2t+1 = row t, column pairs of one parity each),
// which halves the registers a prefetched step takes and doubles the
// wavefronts that hide LDS latency; a lane swap adds the two halves.
// Block J's solution is written over its window values one step later,
// when no thread reads them any more, so one barrier per step suffices.
// Value of the neighbouring lane of a pair (lane ^ 1), by a DPP quad
// permutation [1, 0, 3, 2].
__device__ inline double lane_swap1(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int NB, int BT, int NR, int D, bool FWD>
__device__ inline void band_res_sweep(const double* __restrict__ P, double* const (&vec0)[NR],
                                      const int* __restrict__ xrows, int n, int nblk, int kd,
                                      int ldx, int Jlo = 0, int Jhi = -1, int base = 0) {
  // blocks [Jlo, Jhi) only, updates to rows outside them dropped (a
  // partition's local solve); vec holds rows from `base` on
  if (Jhi < 0) Jhi = nblk;
  const int nst = Jhi - Jlo;
  double* const (&vec)[NR] = vec0;
  constexpr int NH = NB / 2;            // values per thread and step
  const int tid = threadIdx.x;
  const int row = tid >> 1, par = tid & 1;
  double pbuf[D + 1][NH];
  double held[NR];          // this row's block solution, stored a step later
  int held_i = -1;
  // rows past a block's live rows load the last live row's address (the
  // same cache lines, no extra memory traffic) and skip the product
  auto load = [&](int s, double (&pp)[NH]) {
    const int J = FWD ? Jlo + s : Jhi - 1 - s;
    const double2* pj = reinterpret_cast<const double2*>(P + (size_t)J * NB * ldx);
    const int r = min(row, xrows[2 * J + (FWD ? 0 : 1)] - 1);
#pragma unroll
    for (int k = 0; k < NH / 2; ++k) {
      const double2 v = pj[(size_t)(2 * k + par) * ldx + r];
      pp[2 * k] = v.x;
      pp[2 * k + 1] = v.y;
    }
  };
  auto step = [&](int s, double (&pp)[NH]) {
    const int J = FWD ? Jlo + s : Jhi - 1 - s;
    const int j0 = J * NB;
    // Every thread takes delivery of this step's registers here, on every
    // path: consumed only inside the branch below, the loads would count
    // as possibly outstanding at the loop head, where the compiler would
    // then wait for all later prefetches as well.
#pragma unroll
    for (int k = 0; k < NH; ++k) asm volatile("" ::"v"(pp[k]));
    __syncthreads();
    // the previous block's solution (its window is no longer read)
    if (held_i >= 0 && par == 0) {
#pragma unroll
      for (int c = 0; c < NR; ++c) vec[c][held_i - base] = held[c];
    }
    held_i = -1;
    const int rr = row - NB;
    const int i = row < NB ? j0 + row : (FWD ? j0 + NB + rr : j0 - 1 - rr);
    const bool live = row < NB || (row < xrows[2 * J + (FWD ? 0 : 1)] && rr < kd &&
                                   (FWD ? i < Jhi * NB : i >= Jlo * NB));
    // column pair 2k+par of the window: values 4k + 2 par, 4k + 2 par + 1
    double acc[NR][2];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c][0] = acc[c][1] = 0.0;
    if (live) {
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const double2* wv = reinterpret_cast<const double2*>(vec[c] + (j0 - base)) + par;
#pragma unroll
        for (int k0 = 0; k0 < NH / 2; k0 += 4) {
          double2 w[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) w[k] = wv[2 * (k0 + k)];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            acc[c][0] += pp[2 * (k0 + k)] * w[k].x;
            acc[c][1] += pp[2 * (k0 + k) + 1] * w[k].y;
          }
        }
      }
    }
    double v[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const double part = acc[c][0] + acc[c][1];
      v[c] = part + lane_swap1(part);
    }
    if (live && par == 0) {
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        if (row < NB) held[c] = v[c];
        else vec[c][i - base] -= v[c];
      }
      if (row < NB) held_i = i;
    }
  };

  if (nst <= 0) return;
#pragma unroll
  for (int d = 0; d < D; ++d) load(min(d, nst - 1), pbuf[d]);
  // whole groups of D+1 steps with no exit inside a group (an exit between
  // the steps of a group would leave its prefetches unconsumed on a path
  // back to the loop head), then the remaining steps
  int s = 0;
  for (; s + D < nst; s += D + 1) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      const int nx = (u + D) % (D + 1);
      load(min(s + u + D, nst - 1), pbuf[nx]);
      step(s + u, pbuf[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < D; ++u)
    if (s + u < nst) step(s + u, pbuf[u]);
  __syncthreads();
  if (held_i >= 0 && par == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) vec[c][held_i - base] = held[c];
  }
  __syncthreads();
}

// Workgroup-uniform values read from device tables are moved to scalar
// registers: left in vector registers, their loads would stay in the
// in-order memory counter, and every wait for them inside a prefetching
// loop would also wait for all panel loads issued since.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline long long uni64(long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

template <int NB, int NPASS, int BT, int NR, int D>
__global__ __launch_bounds__(BT) void k_band_ext(BandArgs a) {
  extern __shared__ double lds[];
  int md, fsel;
  band_block(a, md, fsel);
  if (md < 0) return;
  const int m = uni(a.dmode[md]);
  // a mode without a second plane sweeps its first plane in that place; a
  // workgroup of a split whose right-hand sides are all such repeats has
  // nothing of its own to do (and must not touch a plane another
  // workgroup is solving)
  if (a.fsplit > 1) {
    bool own = false;
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const int cg = fsel * NR + c;
      own |= a.mode_plane[2 * m + (cg - (cg / a.npl) * a.npl)] >= 0;
    }
    if (!uni(own ? 1 : 0)) return;
  }
  const int n = uni(a.mode_n[m]);
  const long long offp = uni64(a.offP[m]);
  const int nblk = (n + NB - 1) / NB;
  double* ring[NR];
  double* gy[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    const int cg = fsel * NR + c;             // right-hand side across the split
    const int fld = cg / a.npl, kk = cg - fld * a.npl;
    int p = a.mode_plane[2 * m + kk];
    if (p < 0) p = a.mode_plane[2 * m];
    p = uni(p);
    ring[c] = lds + (size_t)c * a.ring;
    gy[c] = a.sol + fld * a.fstride + (size_t)p * a.solstride;
  }
  int* xr = reinterpret_cast<int*>(lds + (size_t)NR * a.ring);
  for (int i = threadIdx.x; i < 2 * nblk; i += BT) xr[i] = a.xrows[(size_t)m * 2 * a.xblk + i];
  __syncthreads();   // the row table is read by the first prefetches
  band_ext_sweep<NB, NPASS, BT, NR, D, true>(a.Pf + offp, ring, gy, xr, n, nblk, a.kd, a.kdp,
                                             a.ring - 1);
  band_ext_sweep<NB, NPASS, BT, NR, D, false>(a.Pb + offp, ring, gy, xr, n, nblk, a.kd, a.kdp,
                                              a.ring - 1);
}

template <int NB, int BT, int NR, int D>
__global__ __launch_bounds__(BT) void k_band_res(BandArgs a) {
  extern __shared__ double lds[];
  int md, fsel;
  band_block(a, md, fsel);
  if (md < 0) return;
  const int m = uni(a.dmode[md]);
  // a mode without a second plane sweeps its first plane in that place; a
  // workgroup of a split whose right-hand sides are all such repeats has
  // nothing of its own to do (and must not touch a plane another
  // workgroup is solving)
  if (a.fsplit > 1) {
    bool own = false;
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const int cg = fsel * NR + c;
      own |= a.mode_plane[2 * m + (cg - (cg / a.npl) * a.npl)] >= 0;
    }
    if (!uni(own ? 1 : 0)) return;
  }
  const int n = uni(a.mode_n[m]);
  const long long offp = uni64(a.offP[m]);
  const int nblk = (n + NB - 1) / NB;
  const int npad = nblk * NB;
  double* vec[NR];
  double* gy[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    const int cg = fsel * NR + c;             // right-hand side across the split
    const int fld = cg / a.npl, kk = cg - fld * a.npl;
    int p = a.mode_plane[2 * m + kk];
    if (p < 0) p = a.mode_plane[2 * m];
    p = uni(p);
    vec[c] = lds + (size_t)c * a.ring;
    gy[c] = a.sol + fld * a.fstride + (size_t)p * a.solstride;
  }
  int* xr = reinterpret_cast<int*>(lds + (size_t)NR * a.ring);
  for (int i = threadIdx.x; i < 2 * nblk; i += BT) xr[i] = a.xrows[(size_t)m * 2 * a.xblk + i];
  for (int i = threadIdx.x; i < npad; i += BT) {
#pragma unroll
    for (int c = 0; c < NR; ++c) vec[c][i] = (i < n) ? gy[c][i] : 0.0;
  }
  __syncthreads();   // the row table is read by the first prefetches
  band_res_sweep<NB, BT, NR, D, true>(a.Pf + offp, vec, xr, n, nblk, a.kd, a.kdp);
  band_res_sweep<NB, BT, NR, D, false>(a.Pb + offp, vec, xr, n, nblk, a.kd, a.kdp);
  for (int i = threadIdx.x; i < n; i += BT) {
#pragma unroll
    for (int c = 0; c < NR; ++c) gy[c][i] = vec[c][i];
  }
}

// ===========================================================================
// Cooperative band solve: W workgroups per mode in a lagged pipeline.
//
// The off-diagonal panel of block J (NB columns x kdp rows) is cut into
// NB-row chunks; chunk c of block J feeds block J+1+c (forward; block
// J-1-c backward).  Lane 0 of a mode (the lead) owns chunks [0, C1): it
// solves each diagonal block, hands y_J to every follower, and applies the
// chunks that feed the next C1 blocks.  Lanes w >= 1 (followers) own
// further chunks; each applies them to y_J into an LDS accumulator per
// target block and hands the finished partial sum of target T to the lead
// once its last contribution to T (from block T-1-c_lo(w)) is in.  The
// lead needs it c_lo(w) >= C1 steps after y_{T-1-c_lo} left, which hides
// the hand-off latency.
//
// Hand-off slots (one per value) hold a signalling-NaN sentinel until the
// producer stores the value; arithmetic never produces that bit pattern, so
// a consumer knows a value has arrived by reading it, with no separate flag
// and no store-completion wait.  Slot stores and loads are device-scope
// relaxed atomics (coherent across XCDs without cache write-back).  Every
// slot is reset to the sentinel by a fill kernel before each launch.  The
// consumers load the slots of step s+PD while working on step s.  Every
// wait is bounded: a timeout sets the error word and the workgroup leaves,
// so the grid always drains.
// ===========================================================================

struct CoopArgs {
  const double* Pf;
  const double* Pb;
  const double* Linv;
  const double* LinvT;
  const int* mode_n;
  const long long* offP;
  const long long* offL;
  const int* dmode;
  const int* mode_plane;
  double* sol;
  double* inbox;            // [dir][md][W-1][nblkmax][NR*NB]  lead -> follower
  double* part;             // [dir][md][W-1][nblkmax][NR*NB]  follower -> lead
                            // (dir stride xstride)
  const int* clo;           // [W+1] chunk ranges per lane
  int* err;
  int solstride, kd, kdp, ndirect, W, nblkmax;
  int npl;                  // planes per field and mode
  long long fstride;        // sol offset between fields
  long long xstride;        // doubles per direction of inbox / part
};

constexpr int COOP_TT = 192;            // panel ("task") threads: the last three waves
constexpr int COOP_RING = 32;           // accumulator blocks per lane (>= chunks + 2)
template <int NB> constexpr int coop_ring() { return NB >= 64 ? 16 : COOP_RING; }
constexpr int COOP_WMAX = 8;            // lanes per mode
constexpr int COOP_PD = 2;              // steps of prefetch
constexpr long COOP_SPIN = 1L << 20;    // polls before a wait gives up
constexpr unsigned long long COOP_SENT = 0x7FF4DEADBEEF0001ull;   // signalling NaN

__device__ inline double ld_dev(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool is_sent(double v) {
  return (unsigned long long)__double_as_longlong(v) == COOP_SENT;
}
// the value of a slot once it has arrived (v: an earlier read of it)
__device__ inline double coop_take(const double* p, double v, int* abort) {
  for (long it = 0; is_sent(v); ++it) {
    if (it >= COOP_SPIN) { *abort = 1; return 0.0; }
    __builtin_amdgcn_s_sleep(1);
    v = ld_dev(p);
  }
  return v;
}

__global__ void k_zero_u32(unsigned* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_coop_fill(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = __longlong_as_double((long long)COOP_SENT);
}

// ===========================================================================
// Partitioned band solve: each mode's band cut into P row ranges at block
// boundaries, one workgroup per (mode, plane, partition), the workgroups of
// a mode on one XCD.  With L = U^T and partition p holding rows [s, e):
//   forward   y_p = z_p - V_p t_{p-1},  z_p = L_p^-1 b_p (the local sweep of
//             the extended panels over the partition's blocks, updates
//             beyond e dropped), t_{p-1} = y[s-kd, s) from partition p-1
//   backward  x_p = w_p - W_p h_{p+1},  w_p = U_p^-1 y_p (local sweep,
//             updates before s dropped), h_{p+1} = x[e, e+kd)
// with the spikes V_p, W_p (n_p x kd, column-major) formed at setup.  The
// tails t and heads h pass along the partitions through hand-off slots
// (the sentinel slots of the cooperative sweep): each workgroup corrects
// its kd boundary rows first and publishes them before correcting the
// rest, so a chain link costs one kd x kd product.  Every wait is bounded;
// a timeout sets the error word.
// ===========================================================================

struct PartArgs {
  BandArgs b;
  const double* spk;          // spikes, all modes
  const long long* spk_off;   // per mode
  const int* prow;            // per mode: P+1 partition row starts
  double* slot;               // [md*F + fsel][part][2][kd]
  int* err;
  int P;
};

// rows [ra, rb) of vec -= S t: S a spike (column-major, ld rows per
// column, ld even, columns 16-byte aligned), t in LDS (kd values).  Row
// pairs with 16-byte loads, eight columns in flight per thread; when the
// rows leave threads idle (the kd boundary rows, on the critical path of
// the chain) the columns are split over up to 8 thread groups whose
// partial sums meet in LDS (red: BT * 2 doubles).
template <int BT>
__device__ inline void spike_apply(double* vec, const double* __restrict__ S, int ld,
                                   const double* t, int kd, int ra, int rb, double* red) {
  const int pa = ra >> 1, pb = (rb + 1) >> 1;      // row pairs [pa, pb)
  const int npair = pb - pa;
  if (npair <= 0) return;
  const int G = max(1, min(8, BT / npair));
  const int tid = threadIdx.x;
  if (G == 1) {
    for (int q = pa + tid; q < pb; q += BT) {
      const double2* col = reinterpret_cast<const double2*>(S + 2 * q);
      const int st = ld >> 1;
      double2 acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
      int c = 0;
      for (; c + 8 <= kd; c += 8) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = col[(size_t)(c + u) * st];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          acc[u & 1].x += v[u].x * t[c + u];
          acc[u & 1].y += v[u].y * t[c + u];
        }
      }
      for (; c < kd; ++c) {
        const double2 v = col[(size_t)c * st];
        acc[0].x += v.x * t[c];
        acc[0].y += v.y * t[c];
      }
      const int r = 2 * q;
      if (r >= ra && r < rb) vec[r] -= acc[0].x + acc[1].x;
      if (r + 1 >= ra && r + 1 < rb) vec[r + 1] -= acc[0].y + acc[1].y;
    }
    return;
  }
  const int q = pa + tid % npair, g = tid / npair;
  if (g < G) {
    const int c0 = (int)((long)g * kd / G), c1 = (int)((long)(g + 1) * kd / G);
    const double2* col = reinterpret_cast<const double2*>(S + 2 * q);
    const int st = ld >> 1;
    double2 acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
    int c = c0;
    for (; c + 8 <= c1; c += 8) {
      double2 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = col[(size_t)(c + u) * st];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        acc[u & 1].x += v[u].x * t[c + u];
        acc[u & 1].y += v[u].y * t[c + u];
      }
    }
    for (; c < c1; ++c) {
      const double2 v = col[(size_t)c * st];
      acc[0].x += v.x * t[c];
      acc[0].y += v.y * t[c];
    }
    red[2 * tid] = acc[0].x + acc[1].x;
    red[2 * tid + 1] = acc[0].y + acc[1].y;
  }
  __syncthreads();
  if (tid < npair) {
    double sx = 0.0, sy = 0.0;
    for (int gg = 0; gg < G; ++gg) {
      sx += red[2 * (gg * npair + tid)];
      sy += red[2 * (gg * npair + tid) + 1];
    }
    const int r = 2 * (pa + tid);
    if (r >= ra && r < rb) vec[r] -= sx;
    if (r + 1 >= ra && r + 1 < rb) vec[r + 1] -= sy;
  }
  __syncthreads();
}

// kd values from a hand-off slot into LDS; false (agreed by the workgroup)
// if the wait timed out
template <int BT>
__device__ inline bool slot_take(const double* slot, double* t, int kd, int* flag) {
  if (threadIdx.x == 0) *flag = 0;
  __syncthreads();
  for (int a = threadIdx.x; a < kd; a += BT) {
    int ab = 0;
    const double v = coop_take(slot + a, ld_dev(slot + a), &ab);
    t[a] = v;
    if (ab) *flag = 1;
  }
  __syncthreads();
  return *flag == 0;
}

template <int NB, int BT>
__global__ __launch_bounds__(BT) void k_band_part(PartArgs pa) {
  extern __shared__ double lds[];
  const BandArgs& a = pa.b;
  const int P = pa.P, F = a.fsplit, G = F * P;
  const int bx = blockIdx.x;
  const int md = (bx / (8 * G)) * 8 + (bx & 7);
  const int rq = (bx >> 3) % G;
  const int fsel = rq / P, part = rq - (rq / P) * P;
  if (md >= a.ndm) return;
  const int m = uni(a.dmode[md]);
  const int fld = fsel / a.npl, kk = fsel - (fsel / a.npl) * a.npl;
  const int pl = a.mode_plane[2 * m + kk];
  if (pl < 0) return;        // every partition of this plane agrees
  double* gy = a.sol + fld * a.fstride + (size_t)uni(pl) * a.solstride;
  const int n = uni(a.mode_n[m]);
  const int nblk = (n + NB - 1) / NB;
  const long long offp = uni64(a.offP[m]);
  const int kd = a.kd;
  const int* pr = pa.prow + (size_t)m * (P + 1);
  const int r0 = uni(pr[part]), r1 = uni(pr[part + 1]);
  const int np_ = r1 - r0;
  const int J0 = r0 / NB, J1 = (part == P - 1) ? nblk : r1 / NB;
  const int nloc = (J1 - J0) * NB;
  double* vec = lds;                                  // a.ring rows
  double* tv = lds + a.ring;                          // kd values (even count)
  double* red = tv + ((kd + 1) / 2) * 2;              // 2 * BT partial sums
  int* flag = reinterpret_cast<int*>(red + 2 * BT);
  int* xr = flag + 2;
  // spikes of this partition (columns of an even number of rows)
  auto ldv = [&](int q) { return ((uni(pr[q + 1]) - uni(pr[q])) + 1) & ~1; };
  const int ld = ldv(part);
  const double* spk = pa.spk + uni64(pa.spk_off[m]);
  const double* V = spk;
  const double* W = spk;
  for (int q = 1; q < P; ++q) {
    const long long sz = (long long)ldv(q) * kd;
    if (q < part) V += sz;
    W += sz;
  }
  for (int q = 0; q < part; ++q) W += (long long)ldv(q) * kd;
  double* sl = pa.slot + (size_t)(md * F + fsel) * P * 2 * kd;
  for (int i = threadIdx.x; i < 2 * nblk; i += BT) xr[i] = a.xrows[(size_t)m * 2 * a.xblk + i];
  for (int i = threadIdx.x; i < nloc; i += BT) vec[i] = (i < np_) ? gy[r0 + i] : 0.0;
  __syncthreads();
  double* const vv[1] = {vec};
  band_res_sweep<NB, BT, 1, 1, true>(a.Pf + offp, vv, xr, n, nblk, kd, a.kdp, J0, J1, r0);
  // forward join: the tail first, published, then the other rows
  bool ok = true;
  if (part > 0) {
    ok = slot_take<BT>(sl + (size_t)(part - 1) * 2 * kd, tv, kd, flag);
    if (ok) {
      spike_apply<BT>(vec, V, ld, tv, kd, np_ - kd, np_, red);
      __syncthreads();
    }
  }
  if (ok && part < P - 1)
    for (int i = threadIdx.x; i < kd; i += BT) st_dev(sl + (size_t)part * 2 * kd + i, vec[np_ - kd + i]);
  if (ok && part > 0) {
    spike_apply<BT>(vec, V, ld, tv, kd, 0, np_ - kd, red);
    __syncthreads();
  }
  if (!ok) {
    if (threadIdx.x == 0) __hip_atomic_store(pa.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  band_res_sweep<NB, BT, 1, 1, false>(a.Pb + offp, vv, xr, n, nblk, kd, a.kdp, J0, J1, r0);
  // backward join: the head first, published, then the other rows
  if (part < P - 1) {
    ok = slot_take<BT>(sl + (size_t)(part + 1) * 2 * kd + kd, tv, kd, flag);
    if (ok) {
      spike_apply<BT>(vec, W, ld, tv, kd, 0, kd, red);
      __syncthreads();
    }
  }
  if (ok && part > 0)
    for (int i = threadIdx.x; i < kd; i += BT) st_dev(sl + (size_t)part * 2 * kd + kd + i, vec[i]);
  if (ok && part < P - 1) {
    spike_apply<BT>(vec, W, ld, tv, kd, kd, np_, red);
    __syncthreads();
  }
  if (!ok) {
    if (threadIdx.x == 0) __hip_atomic_store(pa.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  for (int i = threadIdx.x; i < np_; i += BT) gy[r0 + i] = vec[i];
}

// Threads per workgroup: ceil(NR*NB/64) hand-off waves (one value per
// thread; wave 0 also solves the diagonal block), one store wave, then the
// task waves.  Each group issues its own kind of memory operation (hand-off
// loads, hand-off stores, panel loads), so a wait for one never waits for
// another's (memory counters are per wave and count stores too).
template <int NB, int NR>
constexpr int coop_bt() { return 64 * ((NR * NB + 63) / 64) + 64 + COOP_TT; }

// One sweep; returns false if a hand-off timed out (the workgroup agrees).
template <int NB, int NR, int RPT, bool FWD>
__device__ bool coop_sweep(const CoopArgs& a, int md, int lane, const double* __restrict__ P,
                           const double* __restrict__ L, double* const (&gy)[NR], int n,
                           int nblk, double* acc, double* rhs, double* ybuf, int* abort) {
  constexpr int BT = coop_bt<NB, NR>();
  constexpr int VT = BT - COOP_TT - 64;       // hand-off threads
  constexpr int ST = VT + 64;                 // end of the store wave
  constexpr int QPT = (NB * NB / 64 > 0) ? NB * NB / 64 : 1;   // diagonal: wave 0
  constexpr int G = NB / QPT;
  constexpr int NV = NR * NB;                 // values handed over per block
  const int tid = threadIdx.x;
  const int c_lo = a.clo[lane], c_hi = a.clo[lane + 1];
  const int ntask = (c_hi - c_lo) * NB;
  const int W = a.W;
  // first chunk of every lane, held in registers (a global load inside the
  // step loop would make its wait cover the prefetched hand-off loads)
  int cl[COOP_WMAX];
#pragma unroll
  for (int w = 0; w < COOP_WMAX; ++w) cl[w] = (w < W) ? a.clo[w] : 0x3fffffff;
  const size_t dirofs = FWD ? 0 : (size_t)a.xstride;
  const size_t lstride = (size_t)a.nblkmax * NV;            // per lane
  double* inbox = a.inbox + dirofs + (size_t)md * (W - 1) * lstride;
  double* part = a.part + dirofs + (size_t)md * (W - 1) * lstride;
  const bool lead = lane == 0;
  const int last = lead ? nblk - 1 : nblk - 2 - c_lo;   // last step with work
  constexpr int RING = coop_ring<NB>();
  for (int i = tid; i < RING * NV; i += BT) acc[i] = 0.0;
  if (tid == 0) *abort = 0;

  const bool vth = tid < NV;                  // hand-off threads: one value each
  const int vc = vth ? tid / NB : 0, vi = vth ? tid - (tid / NB) * NB : 0;
  double* gyv = gy[0];                        // this thread's right-hand side
#pragma unroll
  for (int c = 1; c < NR; ++c)
    if (c == vc) gyv = gy[c];
  auto& gyc = gy;
  const bool dact = tid < NB * G;             // within wave 0
  const int da = dact ? tid / G : 0, dg = dact ? tid % G : 0;
  const int tt = tid - ST;                    // task index base (< 0: other waves)
  const bool stw = tid >= VT && tid < ST;     // store wave
  const int sl = tid - VT;                    // lane in the store wave

  // prefetch registers: the panel one step ahead (task waves); hand-off
  // values, b rows and the diagonal-block inverse COOP_PD steps ahead
  // (hand-off waves), indexed by step % COOP_PD
  double pp[RPT][NB];
  double bv[COOP_PD], lv[COOP_PD][QPT], hv[COOP_PD][COOP_WMAX - 1];
  auto block_of = [&](int s) { return FWD ? s : nblk - 1 - s; };
  auto load_panel = [&](int s) {
    if (s > last || ntask <= 0 || tt < 0) return;
    const double* pj = P + (size_t)block_of(s) * NB * a.kdp;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int t = min(tt + q * COOP_TT, ntask - 1);
      const int r = min(c_lo * NB + t, a.kdp - 1);
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) pp[q][jj] = pj[(size_t)jj * a.kdp + r];
    }
  };
  // every hand-off thread issues the same loads on every step (indices
  // clamped, values masked where unused), so the compiler can wait for
  // exactly the loads a step needs and not for the ones issued after them
  auto fetch = [&](int s, int u) {
    if (tid >= VT) return;                    // wave-uniform
    const int sc = min(s, last < 0 ? 0 : last);
    const int J = block_of(sc);
    if (lead) {
      const int row = min(J * NB + vi, n - 1);
      bv[u] = ld_dev(&gyv[row]);
      const double* lj = L + (size_t)J * NB * NB + da * NB + dg * QPT;
#pragma unroll
      for (int q = 0; q < QPT; ++q) lv[u][q] = lj[q];
#pragma unroll
      for (int w = 1; w < COOP_WMAX; ++w)
        hv[u][w - 1] = ld_dev(&part[(min(w, max(W - 1, 1)) - 1) * lstride + (size_t)sc * NV +
                                    min(tid, NV - 1)]);
    } else {
      hv[u][0] = ld_dev(&inbox[(lane - 1) * lstride + (size_t)sc * NV + min(tid, NV - 1)]);
    }
  };

  auto step = [&](int s, int u) -> bool {
    const int J = block_of(s);
    const int j0 = J * NB;
    const int slot = s & (RING - 1);
    if (lead) {
      // A. right-hand side of block J: b - own partials - followers' partials
      if (vth) {
        double v = (j0 + vi < n ? bv[u] : 0.0) - acc[slot * NV + tid];
#pragma unroll
        for (int w = 1; w < COOP_WMAX; ++w)
          if (s >= 1 + cl[w])
            v -= coop_take(&part[(w - 1) * lstride + (size_t)s * NV + tid], hv[u][w - 1], abort);
        rhs[tid] = v;
        acc[slot * NV + tid] = 0.0;
      }
      if (__syncthreads_or(*abort)) return false;
      // B. diagonal block: y_J = L_J rhs (wave 0)
      if (tid < 64) {
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          double accd = 0.0;
#pragma unroll
          for (int q = 0; q < QPT; ++q) accd += lv[u][q] * rhs[c * NB + dg * QPT + q];
          const double v = group_sum<G>(accd);
          if (dact && dg == 0) ybuf[c * NB + da] = v;
        }
      }
      __syncthreads();
      // C. hand y_J to the solution vector and to every follower (store wave)
      if (stw)
        for (int i = sl; i < W * NV; i += 64) {
          const int w = i / NV, k = i - (i / NV) * NV;
          if (w == 0) {
            const int c = k / NB, r = k - (k / NB) * NB;
            if (j0 + r < n) st_dev(&gyc[c][j0 + r], ybuf[k]);
          } else {
            st_dev(&inbox[(w - 1) * lstride + (size_t)s * NV + k], ybuf[k]);
          }
        }
    } else {
      // A. y_J from the lead
      if (vth)
        ybuf[tid] = coop_take(&inbox[(lane - 1) * lstride + (size_t)s * NV + tid], hv[u][0],
                              abort);
      if (__syncthreads_or(*abort)) return false;
    }
    // apply the owned chunks of this block's panel
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int t = tt + q * COOP_TT;
      if (tt >= 0 && t < ntask) {
        const int c = c_lo + t / NB, r = t - (t / NB) * NB;
        const int tslot = (s + 1 + c) & (RING - 1);
        const int tr = FWD ? r : NB - 1 - r;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
          double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
          for (int jj = 0; jj < NB; jj += 4) {
            a0 += pp[q][jj] * ybuf[k * NB + jj];
            a1 += pp[q][jj + 1] * ybuf[k * NB + jj + 1];
            a2 += pp[q][jj + 2] * ybuf[k * NB + jj + 2];
            a3 += pp[q][jj + 3] * ybuf[k * NB + jj + 3];
          }
          acc[(tslot * NR + k) * NB + tr] += (a0 + a1) + (a2 + a3);
        }
      }
    }
    __syncthreads();
    if (!lead) {
      // target s+1+c_lo received its last contribution from this lane
      const int T = s + 1 + c_lo;
      const int tslot = T & (RING - 1);
      if (T < nblk && stw)
        for (int k = sl; k < NV; k += 64) {
          st_dev(&part[(lane - 1) * lstride + (size_t)T * NV + k], acc[tslot * NV + k]);
          acc[tslot * NV + k] = 0.0;
        }
    }
    load_panel(s + 1);
    fetch(s + COOP_PD, u);
    return true;
  };

  __syncthreads();
  load_panel(0);
#pragma unroll
  for (int u = 0; u < COOP_PD; ++u) fetch(u, u);
  for (int s0 = 0; s0 <= last; s0 += COOP_PD) {
#pragma unroll
    for (int u = 0; u < COOP_PD; ++u) {
      if (s0 + u > last) break;
      if (!step(s0 + u, u)) return false;
    }
  }
  return true;
}

template <int NB, int NR, int RPT>
__global__ __launch_bounds__((coop_bt<NB, NR>())) void k_band_coop(CoopArgs a) {
  __shared__ double acc[coop_ring<NB>() * NR * NB];
  __shared__ double rhs[NR * NB];
  __shared__ double ybuf[NR * NB];
  __shared__ int abort_flag;
  // blocks of one mode share an XCD: dispatch is round-robin over 8 XCDs
  const int b = blockIdx.x;
  const int md = (b / (8 * a.W)) * 8 + (b & 7);
  const int lane = (b >> 3) % a.W;
  if (md >= a.ndirect) return;
  const int m = a.dmode[md];
  const int n = a.mode_n[m];
  const int nblk = (n + NB - 1) / NB;
  double* gy[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    const int fld = c / a.npl, kk = c - fld * a.npl;
    int p = a.mode_plane[2 * m + kk];
    if (p < 0) p = a.mode_plane[2 * m];
    gy[c] = a.sol + fld * a.fstride + (size_t)p * a.solstride;
  }
  const bool ok = coop_sweep<NB, NR, RPT, true>(a, md, lane, a.Pf + a.offP[m],
                                                a.Linv + a.offL[m], gy, n, nblk, acc, rhs,
                                                ybuf, &abort_flag);
  if (ok) {
    // the backward sweep re-reads, from global memory, the forward
    // solution that other threads of this workgroup stored
    __threadfence();
    __syncthreads();
    coop_sweep<NB, NR, RPT, false>(a, md, lane, a.Pb + a.offP[m], a.LinvT + a.offL[m], gy, n,
                                   nblk, acc, rhs, ybuf, &abort_flag);
  } else if (threadIdx.x == 0) {
    __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (ok && abort_flag && threadIdx.x == 0)
    __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ===========================================================================
// Cooperative sweep over extended panels.  The panel of block J is cut by
// rows: the lead (lane 0) owns rows [0, NB + C1 NB) -- the block inverse
// and the near chunks -- and each follower w owns the chunks [c_lo(w),
// c_hi(w)) of the update rows (chunk c feeds block J+1+c).  Per step:
//   A  lead: window of block J = b - own partial sums - followers' partial
//      sums (hand-off waves); the store wave writes the previous block's
//      solution out.  Follower: the window of block J from the lead; the
//      store wave hands over the partial sum of block s + c_lo, complete
//      since the previous step.
//   B  lead: the store wave publishes the window to every follower; task
//      waves form the block solution (rows < NB) and the near updates in one
//      product with the window.  Follower: task waves apply their chunks.
// The window goes out before the block is solved, so the followers' work
// is off the lead's critical path.  Hand-off slots, sentinel and bounded
// waits are those of the cooperative sweep above.
// ===========================================================================
template <int NB, int NR, int RPT, bool FWD>
__device__ bool xcoop_sweep(const CoopArgs& a, int md, int lane, const int (&cl)[COOP_WMAX + 1],
                            const double* __restrict__ P, double* const (&gy)[NR], int n,
                            int nblk, double* acc, double* rhs, double* ybuf, int* abort) {
  constexpr int BT = coop_bt<NB, NR>();
  constexpr int VT = BT - COOP_TT - 64;       // hand-off threads
  constexpr int ST = VT + 64;                 // end of the store wave
  constexpr int NV = NR * NB;                 // values handed over per block
  constexpr int RING = coop_ring<NB>();
  const int tid = threadIdx.x;
  const int W = a.W, ldx = a.kdp, kd = a.kd;
  int c_lo = 0, c_hi = cl[1];
#pragma unroll
  for (int w = 1; w < COOP_WMAX; ++w)
    if (w == lane) { c_lo = cl[w]; c_hi = cl[w + 1]; }
  const bool lead = lane == 0;
  const int prow0 = lead ? 0 : NB + c_lo * NB;                      // first owned panel row
  const int nrow = lead ? NB + c_hi * NB : (c_hi - c_lo) * NB;     // owned panel rows
  const size_t dirofs = FWD ? 0 : (size_t)a.xstride;
  const size_t lstride = (size_t)a.nblkmax * NV;            // per lane
  double* inbox = a.inbox + dirofs + (size_t)md * (W - 1) * lstride;
  double* part = a.part + dirofs + (size_t)md * (W - 1) * lstride;
  const int last = lead ? nblk - 1 : nblk - 2 - c_lo;      // last step with work
  for (int i = tid; i < RING * NV; i += BT) acc[i] = 0.0;
  if (tid == 0) *abort = 0;

  const bool vth = tid < NV;                  // hand-off threads: one value each
  const int vc = vth ? tid / NB : 0, vi = vth ? tid - (tid / NB) * NB : 0;
  double* gyv = gy[0];                        // this thread's right-hand side
#pragma unroll
  for (int c = 1; c < NR; ++c)
    if (c == vc) gyv = gy[c];
  auto& gyc = gy;
  const int tt = tid - ST;                    // task index base (< 0: other waves)
  const bool stw = tid >= VT && tid < ST;     // store wave
  const int sl = tid - VT;                    // lane in the store wave

  double pp[COOP_PD][RPT][NB];   // panel rows, COOP_PD steps ahead
  double bv[COOP_PD], hv[COOP_PD][COOP_WMAX - 1];
  auto block_of = [&](int s) { return FWD ? s : nblk - 1 - s; };
  // task waves: the owned rows of one block's panel, 16-byte loads, rows
  // clamped (every task thread issues the same loads on every step)
  auto load_panel = [&](int s, int u) {
    if (tt < 0) return;                       // wave-uniform
    const int sc = min(s, max(last, 0));
    const double2* pj = reinterpret_cast<const double2*>(P + (size_t)block_of(sc) * NB * ldx);
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int t = min(tt + q * COOP_TT, max(nrow - 1, 0));
      const int prow = min(prow0 + t, ldx - 1);
#pragma unroll
      for (int k2 = 0; k2 < NB / 2; ++k2) {
        const double2 v = pj[(size_t)k2 * ldx + prow];
        pp[u][q][2 * k2] = v.x;
        pp[u][q][2 * k2 + 1] = v.y;
      }
    }
  };
  auto fetch = [&](int s, int u) {
    if (tid >= VT) return;                    // wave-uniform
    const int sc = min(s, max(last, 0));
    const int J = block_of(sc);
    if (lead) {
      const int row = min(J * NB + vi, n - 1);
      bv[u] = ld_dev(&gyv[row]);
#pragma unroll
      for (int w = 1; w < COOP_WMAX; ++w)
        hv[u][w - 1] = ld_dev(&part[(min(w, max(W - 1, 1)) - 1) * lstride + (size_t)sc * NV +
                                    min(tid, NV - 1)]);
    } else {
      hv[u][0] = ld_dev(&inbox[(lane - 1) * lstride + (size_t)sc * NV + min(tid, NV - 1)]);
    }
  };
  // the previous block's solution out (lead) / the partial sum of block T
  // to the lead (follower); store wave
  auto flush = [&](int s) {
    if (!stw) return;
    if (lead) {
      const int J = block_of(s);
      for (int k = sl; k < NV; k += 64) {
        const int c = k / NB, r = k - (k / NB) * NB;
        if (J * NB + r < n) st_dev(&gyc[c][J * NB + r], ybuf[k]);
      }
    } else {
      const int T = s + 1 + c_lo;
      if (T < nblk) {
        const int tslot = T & (RING - 1);
        for (int k = sl; k < NV; k += 64) {
          st_dev(&part[(lane - 1) * lstride + (size_t)T * NV + k], acc[tslot * NV + k]);
          acc[tslot * NV + k] = 0.0;
        }
      }
    }
  };

  auto step = [&](int s, int u) -> bool {
    const int J = block_of(s);
    const int j0 = J * NB;
    const int slot = s & (RING - 1);
    // every task thread takes delivery of its panel registers here
    if (tt >= 0) {
#pragma unroll
      for (int q = 0; q < RPT; ++q)
#pragma unroll
        for (int k = 0; k < NB; ++k) asm volatile("" ::"v"(pp[u][q][k]));
    }
    // A
    if (lead) {
      if (vth) {
        double v = (j0 + vi < n ? bv[u] : 0.0) - acc[slot * NV + tid];
#pragma unroll
        for (int w = 1; w < COOP_WMAX; ++w)   // followers only (cl[W] ends the chunks)
          if (w < W && s >= 1 + cl[w])
            v -= coop_take(&part[(w - 1) * lstride + (size_t)s * NV + tid], hv[u][w - 1], abort);
        rhs[tid] = v;
        acc[slot * NV + tid] = 0.0;
      }
    } else if (vth) {
      rhs[tid] = coop_take(&inbox[(lane - 1) * lstride + (size_t)s * NV + tid], hv[u][0], abort);
    }
    if (s > 0) flush(s - 1);
    if (__syncthreads_or(*abort)) return false;
    // B
    if (lead && stw)
      for (int i = sl; i < (W - 1) * NV; i += 64) {
        const int w = i / NV, k = i - (i / NV) * NV;
        st_dev(&inbox[w * lstride + (size_t)s * NV + k], rhs[k]);
      }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int t = tt + q * COOP_TT;
      if (tt >= 0 && t < nrow) {
        const int prow = prow0 + t;
        const int ur = prow - NB;             // update row (< 0: block solution row)
        if (ur < kd) {
          const int c = ur / NB, r = ur - (ur / NB) * NB;
          const int tslot = (s + 1 + c) & (RING - 1);
          const int tr = FWD ? r : NB - 1 - r;
#pragma unroll
          for (int k = 0; k < NR; ++k) {
            const double2* wv = reinterpret_cast<const double2*>(rhs + k * NB);
            double a0 = 0.0, a1 = 0.0;
#pragma unroll
            for (int k2 = 0; k2 < NB / 2; ++k2) {
              const double2 w2 = wv[k2];
              a0 += pp[u][q][2 * k2] * w2.x;
              a1 += pp[u][q][2 * k2 + 1] * w2.y;
            }
            const double v = a0 + a1;
            if (ur < 0) ybuf[k * NB + prow] = v;
            else acc[(tslot * NR + k) * NB + tr] += v;
          }
        }
      }
    }
    __syncthreads();
    load_panel(s + COOP_PD, u);
    fetch(s + COOP_PD, u);
    return true;
  };

  __syncthreads();
#pragma unroll
  for (int u = 0; u < COOP_PD; ++u) {
    load_panel(u, u);
    fetch(u, u);
  }
  // whole groups of COOP_PD steps with no exit inside a group, then the rest
  int s0 = 0;
  for (; s0 + COOP_PD - 1 <= last; s0 += COOP_PD) {
#pragma unroll
    for (int u = 0; u < COOP_PD; ++u)
      if (!step(s0 + u, u)) return false;
  }
#pragma unroll
  for (int u = 0; u < COOP_PD; ++u)
    if (s0 + u <= last && !step(s0 + u, u)) return false;
  if (last >= 0) flush(last);
  __syncthreads();
  return true;
}

template <int NB, int NR, int RPT>
__global__ __launch_bounds__((coop_bt<NB, NR>())) void k_band_xcoop(CoopArgs a) {
  __shared__ double acc[coop_ring<NB>() * NR * NB];
  __shared__ __align__(16) double rhs[NR * NB];   // read as 16-byte pairs
  __shared__ double ybuf[NR * NB];
  __shared__ int abort_flag;
  // blocks of one mode share an XCD: dispatch is round-robin over 8 XCDs
  const int b = blockIdx.x;
  const int md = (b / (8 * a.W)) * 8 + (b & 7);
  const int lane = (b >> 3) % a.W;
  if (md >= a.ndirect) return;
  const int m = uni(a.dmode[md]);
  const int n = uni(a.mode_n[m]);
  const long long offp = uni64(a.offP[m]);
  const int nblk = (n + NB - 1) / NB;
  int cl[COOP_WMAX + 1];
#pragma unroll
  for (int w = 0; w <= COOP_WMAX; ++w) cl[w] = (w <= a.W) ? uni(a.clo[w]) : 0x3fffffff;
  double* gy[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    const int fld = c / a.npl, kk = c - fld * a.npl;
    int p = a.mode_plane[2 * m + kk];
    if (p < 0) p = a.mode_plane[2 * m];
    p = uni(p);
    gy[c] = a.sol + fld * a.fstride + (size_t)p * a.solstride;
  }
  const bool ok = xcoop_sweep<NB, NR, RPT, true>(a, md, lane, cl, a.Pf + offp, gy, n, nblk,
                                                 acc, rhs, ybuf, &abort_flag);
  if (ok) {
    // the backward sweep re-reads the forward solution from global memory
    __threadfence();
    __syncthreads();
    const bool ok2 = xcoop_sweep<NB, NR, RPT, false>(a, md, lane, cl, a.Pb + offp, gy, n, nblk,
                                                     acc, rhs, ybuf, &abort_flag);
    if (!ok2 && threadIdx.x == 0)
      __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (threadIdx.x == 0) {
    __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ __launch_bounds__(ELT_THREADS) void k_backsub(ElemArgs a) {
  extern __shared__ double esh[];
  const int t = threadIdx.x, p = blockIdx.y;
  const int k = a.plane_of[p];
  const int next = a.next, nint = a.nint, npnp = a.npnp;
  const int ce = ELT_THREADS / npnp;
  const int m = k >> 1;
  if (a.mode_kind[m] != SEMX_DIRECT) return;
  const int n = a.mode_n[m];
  double* sub = esh;                 // ce*next
  double* sFi = sub + ce * next;     // ce*nint
  for (int idx = t; idx < ce * npnp; idx += ELT_THREADS) {
    const int le = idx / npnp, q = idx - le * npnp;
    const int ee = blockIdx.x * ce + le;
    if (ee >= a.nel) continue;
    const size_t pe = (size_t)p * a.nel + ee;
    if (q < next) {
      const int gid = a.btog[ee * next + q];
      const double ub = (gid < n) ? a.sol[(size_t)p * a.solstride + gid] : 0.0;
      sub[le * next + q] = ub;
      a.u[(size_t)k * a.planeSize + (size_t)ee * npnp + a.emap[q]] =
          (gid < n) ? ub : (gid >= a.nsolve ? a.g[(size_t)p * a.nzero + (gid - a.nsolve)] : 0.0);
    } else {
      sFi[le * nint + (q - next)] = a.Fi[pe * nint + (q - next)];
    }
  }
  __syncthreads();
  for (int idx = t; idx < ce * nint; idx += ELT_THREADS) {
    const int le = idx / nint, q = idx - le * nint;
    const int ee = blockIdx.x * ce + le;
    if (ee >= a.nel) continue;
    const int set = a.eset[ee];
    const double* hii = a.hii + ((size_t)m * a.nset + set) * nint * nint;
    const double* hbi = a.hbi + ((size_t)m * a.nset + set) * next * nint;
    const double* fi = sFi + le * nint;
    const double* ub = sub + le * next;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int c = 0;
    for (; c + 1 < nint; c += 2) {
      s0 += hii[q * nint + c] * fi[c];
      s1 += hii[q * nint + c + 1] * fi[c + 1];
    }
    if (c < nint) s0 += hii[q * nint + c] * fi[c];
    for (int b = 0; b + 1 < next; b += 2) {
      s2 += hbi[b * nint + q] * ub[b];
      s3 += hbi[(b + 1) * nint + q] * ub[b + 1];
    }
    a.u[(size_t)k * a.planeSize + (size_t)ee * npnp + a.emap[next + q]] = (s0 + s1) - (s2 + s3);
  }
}

// The same for meshes whose elements share one geometric-factor set: one
// block per (direct mode, group of ge elements) with the mode's hii (stored
// transposed) and hbi staged in LDS once.  All of the block's element
// values, both planes, are gathered in one phase (one round of memory
// latency), then each thread forms one interior row of one element for
// both planes, each staged coefficient read serving the two.
__global__ __launch_bounds__(ELT_THREADS) void k_backsub_staged(ElemArgs a,
                                                                const int* __restrict__ dmode,
                                                                const int* __restrict__ mode_plane,
                                                                int ge) {
  extern __shared__ double esh[];
  const int t = threadIdx.x;
  const int next = a.next, nint = a.nint, npnp = a.npnp;
  const int m = dmode[blockIdx.y];
  const int n = a.mode_n[m];
  double* sHiiT = esh;                  // nint*nint, [c][q] = hii[q][c]
  double* sHbi = sHiiT + nint * nint;   // next*nint
  double* sub = sHbi + next * nint;     // 2*ge*next, [plane][element][b]
  double* sFi = sub + 2 * ge * next;    // 2*ge*nint, [plane][element][c]
  {
    const double* hii = a.hii + (size_t)m * nint * nint;
    const double* hbi = a.hbi + (size_t)m * next * nint;
    for (int i = t; i < nint * nint; i += ELT_THREADS) {
      const int q = i / nint, c = i - (i / nint) * nint;
      sHiiT[c * nint + q] = hii[i];
    }
    for (int i = t; i < next * nint; i += ELT_THREADS) sHbi[i] = hbi[i];
  }
  const int e_begin = blockIdx.x * ge, e_end = min(a.nel, e_begin + ge);
  const int ne = e_end - e_begin;
  const int p0 = mode_plane[2 * m], p1 = mode_plane[2 * m + 1];
  const int npl = p1 >= 0 ? 2 : 1;
  for (int idx = t; idx < ne * npnp; idx += ELT_THREADS) {
    const int le = idx / npnp, q = idx - le * npnp;
    const int ee = e_begin + le;
    if (q < next) {
      const int gid = a.btog[ee * next + q];
      for (int pl = 0; pl < npl; ++pl) {
        const int p = pl ? p1 : p0;
        const int k = a.plane_of[p];
        const double ub = (gid < n) ? a.sol[(size_t)p * a.solstride + gid] : 0.0;
        sub[(pl * ge + le) * next + q] = ub;
        a.u[(size_t)k * a.planeSize + (size_t)ee * npnp + a.emap[q]] =
            (gid < n) ? ub
                      : (gid >= a.nsolve ? a.g[(size_t)p * a.nzero + (gid - a.nsolve)] : 0.0);
      }
    } else {
      for (int pl = 0; pl < npl; ++pl) {
        const int p = pl ? p1 : p0;
        sFi[(pl * ge + le) * nint + (q - next)] =
            a.Fi[((size_t)p * a.nel + ee) * nint + (q - next)];
      }
    }
  }
  __syncthreads();
  const int k0 = a.plane_of[p0], k1 = npl > 1 ? a.plane_of[p1] : k0;
  for (int idx = t; idx < ne * nint; idx += ELT_THREADS) {
    const int le = idx / nint, q = idx - le * nint;
    const int ee = e_begin + le;
    const double* fi0 = sFi + le * nint;
    const double* fi1 = sFi + (ge + le) * nint;
    const double* ub0 = sub + le * next;
    const double* ub1 = sub + (ge + le) * next;
    double s0 = 0.0, s1 = 0.0, t0 = 0.0, t1 = 0.0;
    for (int c = 0; c < nint; ++c) {
      const double h = sHiiT[c * nint + q];
      s0 += h * fi0[c];
      s1 += h * fi1[c];
    }
    for (int b = 0; b < next; ++b) {
      const double h = sHbi[b * nint + q];
      t0 += h * ub0[b];
      t1 += h * ub1[b];
    }
    const size_t off = (size_t)ee * npnp + a.emap[next + q];
    a.u[(size_t)k0 * a.planeSize + off] = s0 - t0;
    if (npl > 1) a.u[(size_t)k1 * a.planeSize + off] = s1 - t1;
  }
}

// ===========================================================================
// Jacobi-preconditioned CG, persistent workgroups (one per plane, or W
// sharing a plane, below): semtex's JACPCG branch of Field::solve (src/field.cpp:627-720) with the global
// operator of Field::HelmholtzOperator (src/field.cpp:771-856).
// Vectors are globally numbered: nglobal element-boundary nodes, then the
// element interiors in emap order.  Rows of essential (and pinned) nodes are
// held at zero.  T is the precision of the iteration vectors; with
// refinement the FP64 residual b - A x is recomputed after each inner T
// solve and the correction added to the FP64 solution.
// ===========================================================================

struct PcgArgs {
  int np, npnp, next, nint, nel, nglobal, nsolve, nzero, npts, planeSize, nset;
  double lambda2, beta;
  int base;
  const double* hcon;     // per mode: lambda2 + betak2 (1 with mq)
  const double* mq;       // cylindrical: per mode and set, operator mass factors
  const double* G;
  const int* eset;
  const int* emap;
  const int* pmap;
  const int* btog;
  const double* ops;
  const int* asm_ptr;
  const int* asm_idx;
  const int* plane_of;
  const int* pcg_plane;   // PCG plane-list indices
  const int* mode_n;
  const double* pc;       // nmodes*npts
  const float* pc32;      // nmodes*npts
  const double* Fb;
  const double* Fi;
  const double* g;
  double* u;
  double* work;           // per PCG plane, workstride doubles
  size_t workstride;
  double tol;             // TOL_REL
  int step_max;           // STEP_MAX (per inner solve with refinement)
  double inner_tol;       // inner relative tolerance with refinement
  int max_outer;
  int* iters;             // per PCG plane: iterations done (negative: limit hit)
  int W;                  // workgroups per plane (without wmap)
  const int* wmap;        // or per workgroup: plane slot, part, parts, 0
  unsigned* pctr;         // per plane: barrier arrivals (zero at launch)
  double* pslot;          // per plane: 2 x PCG_WMAX x 2 partial sums
  int* err;               // set when a barrier wait timed out
};

constexpr int PCG_THREADS = 1024;
constexpr int PCG_WMAX = 16;            // workgroups of one PCG plane at most
constexpr int PCG_VU = 4;     // vector-pass entries per thread per load group
constexpr int PCG_TUNE_ITERS = 30;   // iterations per timed operator candidate

// Block barrier of the PCG kernels, whose vectors live in global memory and
// are written and read by different threads of the workgroup: this
// wavefront's outstanding vector-memory operations are waited for first.
// The workgroup barrier as compiled for gfx950 waits for LDS operations
// only, and a global store still in flight when it releases can be missed
// by a wavefront that loads the same address after it (seen as results
// that differ from run to run).
__device__ __forceinline__ void pcg_barrier() {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

// The PCG kernels' LDS reduction scratch `red`: two buffers of 2 values per
// wavefront, used in turn by consecutive block sums (so one barrier per sum
// suffices: the barrier that ends a sum also separates the readers of the
// previous sum's buffer from the writers of the sum after it), then the
// grid-barrier timeout flag.
constexpr int RED_BUF = 2 * (PCG_THREADS / 64);
constexpr int RED_FLAG = 2 * RED_BUF;
constexpr int RED_DOUBLES = RED_FLAG + 8;

// Sums of v and u over the workgroup's threads (each added in the same
// order as a sum of one value: wavefront shuffles, then the wavefronts in
// order).  par: the buffer to use, flipped on return; uniform.
__device__ __forceinline__ double2 block_sum2(double v, double u, double* red, int& par) {
#pragma unroll
  for (int off = 32; off >= 1; off /= 2) {
    v += __shfl_xor(v, off);
    u += __shfl_xor(u, off);
  }
  const int wave = threadIdx.x >> 6;
  double* rb = red + par * RED_BUF;
  par ^= 1;
  if ((threadIdx.x & 63) == 0) {
    rb[2 * wave] = v;
    rb[2 * wave + 1] = u;
  }
  pcg_barrier();
  double2 t = make_double2(0.0, 0.0);
#pragma unroll
  for (int w = 0; w < PCG_THREADS / 64; ++w) {
    t.x += rb[2 * w];
    t.y += rb[2 * w + 1];
  }
  return t;
}

// Workgroups sharing one plane's iteration.  W = 1: the whole plane in one
// workgroup (every range below is the full one, part_sync is a block
// barrier).  W > 1: workgroup w owns the vector rows [vlo, vhi), the
// elements [elo, ehi) of the operator and the global rows [glo, ghi) of its
// assembly; the W workgroups meet at a grid barrier (an arrival counter in
// device memory, agent-scope fences so that their vector writes are seen
// across XCDs) before every gather of another workgroup's data, and dot
// products add the W workgroups' partial sums in workgroup order, so all
// of them take the same value.  Every wait is bounded: a timeout sets the
// error word and the workgroup stops waiting (its results are then wrong
// and the error is reported by the next check).
struct PcgPart {
  int W = 1, w = 0;
  int vlo = 0, vhi = 0;       // vector rows
  int elo = 0, ehi = 0;       // elements of the operator
  int glo = 0, ghi = 0;       // assembled global rows
  unsigned* ctr = nullptr;    // arrivals at this plane's barriers
  double* slot = nullptr;     // [2 parities][W][2] partial sums
  int* err = nullptr;
  unsigned epoch = 0;         // barriers passed
  int rpar = 0;               // reduction buffer of the next block sum
};

__device__ inline PcgPart part_of(int W, int w, int npts, int nel, int nglobal, unsigned* ctr,
                                  double* slot, int* err) {
  PcgPart P;
  P.W = W; P.w = w;
  P.vlo = (int)((long long)npts * w / W); P.vhi = (int)((long long)npts * (w + 1) / W);
  P.elo = (int)((long long)nel * w / W); P.ehi = (int)((long long)nel * (w + 1) / W);
  P.glo = (int)((long long)nglobal * w / W); P.ghi = (int)((long long)nglobal * (w + 1) / W);
  P.ctr = ctr; P.slot = slot; P.err = err;
  // workgroup-uniform: kept in scalar registers
  P.W = uni(P.W); P.w = uni(P.w);
  P.vlo = uni(P.vlo); P.vhi = uni(P.vhi); P.elo = uni(P.elo); P.ehi = uni(P.ehi);
  P.glo = uni(P.glo); P.ghi = uni(P.ghi);
  return P;
}

constexpr long PCG_SPIN = 1L << 22;    // polls before a barrier wait gives up

// A barrier of the plane's workgroups (a block barrier when W = 1).  red:
// the workgroup's reduction scratch (entry RED_FLAG carries a timeout to every
// thread).
__device__ __forceinline__ void part_sync(PcgPart& P, double* red) {
  pcg_barrier();
  if (P.W == 1) return;
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(P.ctr, 1u);
    const unsigned target = (P.epoch + 1) * (unsigned)P.W;
    long it = 0;
    bool late = false;
    while (__hip_atomic_load(P.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++it > PCG_SPIN) { late = true; break; }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
    if (late) __hip_atomic_store(P.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    red[RED_FLAG] = late ? 1.0 : 0.0;
  }
  pcg_barrier();
  P.epoch += 1;
  if (red[RED_FLAG] != 0.0) P.W = 1;    // stop waiting (the error word is set)
}

// sums of v and u over the threads of every workgroup of the plane
__device__ __forceinline__ double2 psum2(double v, double u, double* red, PcgPart& P) {
  const double2 ab = block_sum2(v, u, red, P.rpar);
  const double a = ab.x, b = ab.y;
  if (P.W == 1) return make_double2(a, b);
  double* sl = P.slot + (size_t)(P.epoch & 1) * P.W * 2;
  if (threadIdx.x == 0) {
    __hip_atomic_store(sl + 2 * P.w, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(sl + 2 * P.w + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const int W = P.W;
  part_sync(P, red);
  if (P.W == 1) return make_double2(a, b);   // timed out
  double x = 0.0, y = 0.0;
  for (int k = 0; k < W; ++k) {
    x += __hip_atomic_load(sl + 2 * k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    y += __hip_atomic_load(sl + 2 * k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return make_double2(x, y);
}

__device__ __forceinline__ double psum(double v, double* red, PcgPart& P) {
  if (P.W == 1) return block_sum2(v, 0.0, red, P.rpar).x;
  return psum2(v, 0.0, red, P).x;
}

// What the matrix-free operator of one plane needs.
struct PcgOp {
  int np, npnp, next, nint, nel, nglobal, npts;
  double hcon;            // lambda2 + betak2
  bool svv;               // SVV derivative operators (lambda2 > EPSDP)
  const double* G;
  const int* eset;
  const int* pmap;
  const int* btog;
  const double* ops;
  const int* asm_ptr;
  const int* asm_idx;
  int nset;               // geometric-factor sets (1: the same for all elements)
  const double* mq;       // cylindrical: this mode's mass factors per set
                          // (nset*npnp, hcon = 1); null: Q4 of the set
};

// The mass factors of the operator for the elements of set `set`.
__device__ __forceinline__ const double* pcg_mass(const PcgOp& a, int set, const double* g) {
  return a.mq ? a.mq + (size_t)set * a.npnp : g + G_Q4 * a.npnp;
}

// Owner-gather assembly of the element-boundary results lb into rows
// [0, nglobal) of y (Field::HelmholtzOperator's local2globalSum,
// src/field.cpp:771-856, without atomics), returning this thread's share
// of w . y.  PCG_VU rows per thread at a time, the index loads of all of
// them (up to four contributions each; more are summed after, in order)
// issued before the value loads: the sums are the same as one row at a time.
template <typename T, bool MASK>
__device__ __forceinline__ double pcg_assemble(const PcgOp& a, int n, const T* __restrict__ lb,
                                      T* __restrict__ y, const T* __restrict__ w,
                                      const PcgPart& P) {
  const int tid = threadIdx.x, ng = a.nglobal;
  double wy = 0.0;
  for (int g0 = P.glo + tid; g0 < P.ghi; g0 += PCG_THREADS * PCG_VU) {
    int c0[PCG_VU], c1[PCG_VU], ix[PCG_VU][4];
#pragma unroll
    for (int u = 0; u < PCG_VU; ++u) {
      const int gI = g0 + u * PCG_THREADS;
      const bool live = gI < P.ghi && (!MASK || gI < n);
      const int gc = min(gI, ng - 1);
      c0[u] = a.asm_ptr[gc];
      c1[u] = live ? a.asm_ptr[gc + 1] : c0[u];
    }
#pragma unroll
    for (int u = 0; u < PCG_VU; ++u)
#pragma unroll
      for (int k = 0; k < 4; ++k) ix[u][k] = (c0[u] + k < c1[u]) ? a.asm_idx[c0[u] + k] : -1;
    T lv[PCG_VU][4];
#pragma unroll
    for (int u = 0; u < PCG_VU; ++u)
#pragma unroll
      for (int k = 0; k < 4; ++k) lv[u][k] = (ix[u][k] >= 0) ? lb[ix[u][k]] : (T)0;
#pragma unroll
    for (int u = 0; u < PCG_VU; ++u) {
      const int gI = g0 + u * PCG_THREADS;
      T v = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (ix[u][k] >= 0) v += lv[u][k];
      for (int c = c0[u] + 4; c < c1[u]; ++c) v += lb[a.asm_idx[c]];
      if (gI < P.ghi) {
        y[gI] = v;
        wy += (double)w[gI] * (double)v;
      }
    }
  }
  return wy;
}

// y = A x on the globally numbered vector; returns the dot product w . y
// (w may be x itself).  MASK: unknowns >= n are treated as zero and rows
// [n, nglobal) of y are zero (the solved system); otherwise every node takes
// part (the lifting of essential values).
template <typename T, bool MASK = true>
__device__ __forceinline__ double pcg_matvec(const PcgOp& a, int n, const T* __restrict__ x,
                             T* __restrict__ y, T* __restrict__ lb, const T* __restrict__ w,
                             T* sh, double* red, PcgPart& P) {
  const int tid = threadIdx.x, np = a.np, npnp = a.npnp, next = a.next, nint = a.nint;
  const int ce = PCG_THREADS / npnp;
  const T hcon = (T)a.hcon;
  const bool svv = a.svv;
  T* sD = sh;
  T* sT = sD + npnp;
  T* sU = sT + npnp;
  T* sR = sU + ce * npnp;
  T* sS = sR + ce * npnp;
  if (tid < npnp) {
    sD[tid] = (T)a.ops[(svv ? 2 : 0) * npnp + tid];
    sT[tid] = (T)a.ops[(svv ? 3 : 1) * npnp + tid];
  }
  const int el = tid / npnp, pt = tid - (tid / npnp) * npnp;
  const int i = pt / np, j = pt - (pt / np) * np;
  const int pm = a.pmap[min(pt, npnp - 1)];
  // x was just written by other threads (the caller's vector pass): every
  // gather below comes after this barrier
  part_sync(P, red);
  // gather pipeline: the global index of this thread's node two passes
  // ahead, its value one pass ahead (-1: masked, value zero)
  auto gidx = [&](int e0) {
    const int e = min(e0 + el, a.nel - 1);
    if (pm < next) {
      const int gid = a.btog[e * next + pm];
      return (!MASK || gid < n) ? gid : -1;
    }
    return a.nglobal + e * nint + (pm - next);
  };
  int ia = gidx(P.elo), ib = gidx(P.elo + ce);
  T vnext = ia >= 0 ? x[ia] : (T)0;
  double wy = 0.0;
  for (int e0 = P.elo; e0 < P.ehi; e0 += ce) {
    const int e = e0 + el;
    const bool act = el < ce && e < P.ehi;
    const int base = el * npnp;
    const T v = vnext;
    pcg_barrier();
    if (act) sU[base + pt] = v;
    vnext = ib >= 0 ? x[ib] : (T)0;
    ib = gidx(e0 + 2 * ce);
    pcg_barrier();
    const int set = act ? a.eset[e] : 0;
    const double* g = a.G + (size_t)set * G_N * npnp;
    if (act) {
      T r = 0, s2 = 0;
      for (int q = 0; q < np; ++q) {
        r += sU[base + i * np + q] * sT[q * np + j];
        s2 += sD[i * np + q] * sU[base + q * np + j];
      }
      const T q1 = (T)g[G_Q1 * npnp + pt], q2 = (T)g[G_Q2 * npnp + pt];
      const T q3 = (T)g[G_Q3 * npnp + pt];
      sR[base + pt] = q1 * r + q3 * s2;
      sS[base + pt] = q2 * s2 + q3 * r;
    }
    pcg_barrier();
    if (act) {
      T out = (T)pcg_mass(a, set, g)[pt] * sU[base + pt] * hcon;
      T s1 = 0, s3 = 0;
      for (int q = 0; q < np; ++q) {
        s1 += sT[i * np + q] * sS[base + q * np + j];
        s3 += sR[base + i * np + q] * sD[q * np + j];
      }
      out += s1;
      out += s3;
      if (pm < next) {
        lb[e * next + pm] = out;
      } else {
        const int idx = a.nglobal + e * nint + (pm - next);
        y[idx] = out;
        wy += (double)w[idx] * (double)out;
      }
    }
  }
  part_sync(P, red);
  wy += pcg_assemble<T, MASK>(a, n, lb, y, w, P);
  return psum(wy, red, P);
}

// PCG iterations on A x = r (r holds the residual of the initial x and is
// updated in place) until r.r <= epsb2 or the step limit; semtex's loop.
// The same operator with each wavefront working through its own elements
// (e = wave, wave + 16, ...) in a private LDS slice with wavefront-level
// synchronisation only, the next element's values already in registers
// while the current one is computed; one block barrier then separates the
// element phase from the owner-gather assembly.  Fewer block barriers, at
// the cost of idle lanes when np^2 is not a multiple of 64: faster for the
// small, many-iteration problems of the batch kernel, slower for one large
// system per workgroup.  LDS: (2 + 3 PCG_WAVES) np^2 values of T.
constexpr int PCG_WAVES = PCG_THREADS / 64;
constexpr int PCG_NPT = 3;                    // nodes per lane (np <= 13)

template <typename T, bool MASK = true>
__device__ __forceinline__ double pcg_matvec_wave(const PcgOp& a, int n, const T* __restrict__ x,
                                  T* __restrict__ y, T* __restrict__ lb,
                                  const T* __restrict__ w, T* sh, double* red, PcgPart& P) {
  const int tid = threadIdx.x, np = a.np, npnp = a.npnp, next = a.next, nint = a.nint;
  const int lane = tid & 63, wave = tid >> 6;
  const T hcon = (T)a.hcon;
  const bool svv = a.svv;
  T* sD = sh;
  T* sT = sD + npnp;
  T* sU = sT + npnp + wave * 3 * npnp;
  T* sR = sU + npnp;
  T* sS = sR + npnp;
  part_sync(P, red);   // x is complete, the previous use of the operators too
  if (tid < npnp) {
    sD[tid] = (T)a.ops[(svv ? 2 : 0) * npnp + tid];
    sT[tid] = (T)a.ops[(svv ? 3 : 1) * npnp + tid];
  }
  pcg_barrier();
  // this lane's nodes: pt = lane + 64 k; gather pipeline: global indices
  // two elements ahead, values one ahead (-1: masked, value zero)
  int pmv[PCG_NPT];
#pragma unroll
  for (int k = 0; k < PCG_NPT; ++k) pmv[k] = a.pmap[min(lane + 64 * k, npnp - 1)];
  auto gidx = [&](int e, int (&g)[PCG_NPT]) {
    const int ec = min(e, a.nel - 1);
#pragma unroll
    for (int k = 0; k < PCG_NPT; ++k) {
      if (pmv[k] < next) {
        const int gid = a.btog[ec * next + pmv[k]];
        g[k] = (!MASK || gid < n) ? gid : -1;
      } else {
        g[k] = a.nglobal + ec * nint + (pmv[k] - next);
      }
    }
  };
  auto gval = [&](const int (&g)[PCG_NPT], T (&v)[PCG_NPT]) {
#pragma unroll
    for (int k = 0; k < PCG_NPT; ++k) v[k] = g[k] >= 0 ? x[g[k]] : (T)0;
  };
  double wy = 0.0;
  T cur[PCG_NPT], nxt[PCG_NPT];
  int gi[PCG_NPT];
  gidx(P.elo + wave, gi);
  gval(gi, cur);
  gidx(P.elo + wave + PCG_WAVES, gi);
  for (int e = P.elo + wave; e < P.ehi; e += PCG_WAVES) {
    gval(gi, nxt);
    gidx(e + 2 * PCG_WAVES, gi);
#pragma unroll
    for (int k = 0; k < PCG_NPT; ++k)
      if (lane + 64 * k < npnp) sU[lane + 64 * k] = cur[k];
    wave_sync();
    const double* g = a.G + (size_t)a.eset[e] * G_N * npnp;
#pragma unroll 1
    for (int k = 0; k < PCG_NPT; ++k) {
      const int pt = lane + 64 * k;
      if (pt < npnp) {
        const int i = pt / np, j = pt - (pt / np) * np;
        T r = 0, s2 = 0;
        for (int q = 0; q < np; ++q) {
          r += sU[i * np + q] * sT[q * np + j];
          s2 += sD[i * np + q] * sU[q * np + j];
        }
        const T q1 = (T)g[G_Q1 * npnp + pt], q2 = (T)g[G_Q2 * npnp + pt];
        const T q3 = (T)g[G_Q3 * npnp + pt];
        sR[pt] = q1 * r + q3 * s2;
        sS[pt] = q2 * s2 + q3 * r;
      }
    }
    wave_sync();
#pragma unroll 1
    for (int k = 0; k < PCG_NPT; ++k) {
      const int pt = lane + 64 * k;
      if (pt < npnp) {
        const int i = pt / np, j = pt - (pt / np) * np;
        const int pmk = a.pmap[pt];
        T out = (T)pcg_mass(a, a.eset[e], g)[pt] * sU[pt] * hcon;
        T s1 = 0, s3 = 0;
        for (int q = 0; q < np; ++q) {
          s1 += sT[i * np + q] * sS[q * np + j];
          s3 += sR[i * np + q] * sD[q * np + j];
        }
        out += s1;
        out += s3;
        if (pmk < next) {
          lb[e * next + pmk] = out;
        } else {
          const int idx = a.nglobal + e * nint + (pmk - next);
          y[idx] = out;
          wy += (double)w[idx] * (double)out;
        }
      }
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < PCG_NPT; ++k) cur[k] = nxt[k];
  }
  part_sync(P, red);
  wy += pcg_assemble<T, MASK>(a, n, lb, y, w, P);
  return psum(wy, red, P);
}

// The operator on the BF16 matrix cores (SEMX_PCG_BF16_REFINED): each
// wavefront takes one element at a time, zero-padded to 16 x 16, and forms
// the four tensor-product contractions of HelmholtzOp (src/element.cpp:
// 1621-1650; R = U T, S = D U, then T S' + R' D with R' = Q1 R + Q3 S,
// S' = Q2 S + Q3 R) as four v_mfma_f32_16x16x16_bf16 with FP32
// accumulation: operands rounded to BF16 (nearest even), everything else --
// the geometric factors, the mass term (lambda2 + betak2) Q4 U and the
// sums -- in FP32.  Lane l holds, for v = 0..3, A operands at (row l%16,
// column 4(l/16)+v), B operands and results at (row 4(l/16)+v, column
// l%16); U and R' change from the second layout to the first through a
// private LDS tile (stride 20 floats) with wavefront synchronisation only.
// The next element's values are gathered while the current one is formed.
// np <= 16; LDS: PCG_WAVES * 2 * 16 * 20 floats.
typedef short mf_bf4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
constexpr int MF_LD = 20;                   // LDS tile row stride (floats)
constexpr int MF_TILE = 16 * MF_LD;

__device__ inline short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ inline mf_bf4 bf16x4(float a, float b, float c, float d) {
  mf_bf4 r;
  r[0] = bf16_rne(a); r[1] = bf16_rne(b); r[2] = bf16_rne(c); r[3] = bf16_rne(d);
  return r;
}

template <bool MASK = true>
__device__ __forceinline__ double pcg_matvec_mfma(const PcgOp& a, int n, const float* __restrict__ x,
                                  float* __restrict__ y, float* __restrict__ lb,
                                  const float* __restrict__ w, float* sh, double* red, PcgPart& P) {
  const int tid = threadIdx.x, np = a.np, npnp = a.npnp, next = a.next, nint = a.nint;
  const int lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, rb = 4 * (lane >> 4);
  float* tU = sh + wave * 2 * MF_TILE;
  float* tR = tU + MF_TILE;
  const double* D = a.ops + (a.svv ? 2 : 0) * npnp;
  const double* Tm = a.ops + (a.svv ? 3 : 1) * npnp;
  // operator operands of this lane, zero outside np x np
  auto ent = [&](const double* M, int i, int j) {
    return (i < np && j < np) ? (float)M[i * np + j] : 0.0f;
  };
  const mf_bf4 aD = bf16x4(ent(D, c, rb), ent(D, c, rb + 1), ent(D, c, rb + 2), ent(D, c, rb + 3));
  const mf_bf4 aT = bf16x4(ent(Tm, c, rb), ent(Tm, c, rb + 1), ent(Tm, c, rb + 2),
                           ent(Tm, c, rb + 3));
  const mf_bf4 bT = bf16x4(ent(Tm, rb, c), ent(Tm, rb + 1, c), ent(Tm, rb + 2, c),
                           ent(Tm, rb + 3, c));
  const mf_bf4 bD = bf16x4(ent(D, rb, c), ent(D, rb + 1, c), ent(D, rb + 2, c), ent(D, rb + 3, c));
  // this lane's four nodes (row rb+v, column c): emap index or -1
  int pm[4], pt[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const bool in = rb + v < np && c < np;
    pt[v] = in ? (rb + v) * np + c : 0;
    pm[v] = in ? a.pmap[pt[v]] : -1;
  }
  const float hcon = (float)a.hcon;
  // gather pipeline: global indices two elements ahead, values one ahead
  // (-1: padding or masked, value zero)
  auto gidx = [&](int e, int (&g)[4]) {
    const int ec = min(e, a.nel - 1);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int gid = -1;
      if (pm[v] >= 0) {
        if (pm[v] < next) {
          gid = a.btog[ec * next + pm[v]];
          if (MASK && gid >= n) gid = -1;
        } else {
          gid = a.nglobal + ec * nint + (pm[v] - next);
        }
      }
      g[v] = gid;
    }
  };
  auto gval = [&](const int (&g)[4], float (&u)[4]) {
#pragma unroll
    for (int v = 0; v < 4; ++v) u[v] = g[v] >= 0 ? x[g[v]] : 0.0f;
  };
  double wy = 0.0;
  float cur[4], nxt[4];
  int gi[4];
  // x was just written by other threads (the caller's vector pass): every
  // gather comes after this barrier
  part_sync(P, red);
  gidx(P.elo + wave, gi);
  gval(gi, cur);
  gidx(P.elo + wave + PCG_WAVES, gi);
  // geometric factors of this lane's nodes: loaded once when every element
  // shares one set, per element otherwise
  float q1[4], q2[4], q3[4], q4[4];
  auto qload = [&](int e) {
    const double* g = a.G + (size_t)a.eset[e] * G_N * npnp;
    const double* gm = pcg_mass(a, a.eset[e], g);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool in = pm[v] >= 0;
      q1[v] = in ? (float)g[G_Q1 * npnp + pt[v]] : 0.0f;
      q2[v] = in ? (float)g[G_Q2 * npnp + pt[v]] : 0.0f;
      q3[v] = in ? (float)g[G_Q3 * npnp + pt[v]] : 0.0f;
      q4[v] = in ? (float)gm[pt[v]] : 0.0f;
    }
  };
  const bool one_set = a.nset == 1;
  if (one_set) qload(0);
  for (int e = P.elo + wave; e < P.ehi; e += PCG_WAVES) {
    gval(gi, nxt);
    gidx(e + 2 * PCG_WAVES, gi);
    if (!one_set) qload(e);
#pragma unroll
    for (int v = 0; v < 4; ++v) tU[(rb + v) * MF_LD + c] = cur[v];
    wave_sync();
    const mf_f4 ua = *(const mf_f4*)&tU[c * MF_LD + rb];
    const mf_f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    const mf_f4 S = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
        aD, bf16x4(cur[0], cur[1], cur[2], cur[3]), zero, 0, 0, 0);
    const mf_f4 R = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
        bf16x4(ua[0], ua[1], ua[2], ua[3]), bT, zero, 0, 0, 0);
    float rp[4], sp[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      rp[v] = q1[v] * R[v] + q3[v] * S[v];
      sp[v] = q2[v] * S[v] + q3[v] * R[v];
      tR[(rb + v) * MF_LD + c] = rp[v];
    }
    wave_sync();
    const mf_f4 ra = *(const mf_f4*)&tR[c * MF_LD + rb];
    mf_f4 o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aT, bf16x4(sp[0], sp[1], sp[2], sp[3]),
                                                        zero, 0, 0, 0);
    o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16x4(ra[0], ra[1], ra[2], ra[3]), bD, o, 0,
                                                  0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (pm[v] < 0) continue;
      const float out = o[v] + q4[v] * cur[v] * hcon;
      if (pm[v] < next) {
        lb[e * next + pm[v]] = out;
      } else {
        const int idx = a.nglobal + e * nint + (pm[v] - next);
        y[idx] = out;
        wy += (double)w[idx] * (double)out;
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) cur[v] = nxt[v];
  }
  part_sync(P, red);
  wy += pcg_assemble<float, MASK>(a, n, lb, y, w, P);
  return psum(wy, red, P);
}

// Element-operator variants: the block-wide and wavefront-level scalar
// operators (any T) and the BF16 matrix-core operator (T = float).
enum { OPK_BLOCK = 0, OPK_WAVE = 1, OPK_MFMA = 2 };

template <typename T, bool MASK, int OPK>
__device__ __forceinline__ double pcg_op(const PcgOp& a, int n, const T* x, T* y, T* lb, const T* w,
                                T* sh, double* red, PcgPart& P) {
  if constexpr (OPK == OPK_MFMA) {
    static_assert(std::is_same<T, float>::value, "the BF16 operator works on FP32 vectors");
    return pcg_matvec_mfma<MASK>(a, n, x, y, lb, w, sh, red, P);
  } else if constexpr (OPK == OPK_WAVE) {
    return pcg_matvec_wave<T, MASK>(a, n, x, y, lb, w, sh, red, P);
  } else {
    return pcg_matvec<T, MASK>(a, n, x, y, lb, w, sh, red, P);
  }
}

template <typename T, int OPK = OPK_BLOCK>
__device__ __forceinline__ int pcg_core(const PcgOp& a, int n, T* x, T* r, T* p, T* q, T* lb, const T* pc,
                        double epsb2, int step_max, T* sh, double* red, PcgPart& P) {
  const int tid = threadIdx.x, npts = a.npts;
  double r2 = 0.0, rho1 = 0.0, rho2 = 1.0;
  for (int i = P.vlo + tid; i < P.vhi; i += PCG_THREADS) {
    const double ri = (double)r[i];
    r2 += ri * ri;
    rho1 += ri * (double)pc[i] * ri;
  }
  {
    const double2 t = psum2(r2, rho1, red, P);
    r2 = t.x;
    rho1 = t.y;
  }
  int it = 0;
  // The vector passes take PCG_VU entries per thread at a time (i = tid +
  // (g PCG_VU + u) PCG_THREADS, the same order as one at a time, so the
  // sums are unchanged) with every load issued before the first use: one
  // memory latency per group instead of one per entry.
  while (r2 > epsb2 && ++it < step_max) {
    const T beta = (it == 1) ? (T)0 : (T)(rho1 / rho2);
    for (int i0 = P.vlo + tid; i0 < P.vhi; i0 += PCG_THREADS * PCG_VU) {
      T cv[PCG_VU], rv[PCG_VU], pv[PCG_VU];
#pragma unroll
      for (int u = 0; u < PCG_VU; ++u) {
        const int i = min(i0 + u * PCG_THREADS, P.vhi - 1);
        cv[u] = pc[i];
        rv[u] = r[i];
        pv[u] = p[i];
      }
#pragma unroll
      for (int u = 0; u < PCG_VU; ++u) {
        const int i = i0 + u * PCG_THREADS;
        const T z = cv[u] * rv[u];
        if (i < P.vhi) p[i] = (it == 1) ? z : z + beta * pv[u];
      }
    }
    const double pq = pcg_op<T, true, OPK>(a, n, p, q, lb, p, sh, red, P);
    const T alpha = (T)(rho1 / pq);
    double nr2 = 0.0, nrho = 0.0;
    for (int i0 = P.vlo + tid; i0 < P.vhi; i0 += PCG_THREADS * PCG_VU) {
      T xv[PCG_VU], pv[PCG_VU], rv[PCG_VU], qv[PCG_VU], cv[PCG_VU];
#pragma unroll
      for (int u = 0; u < PCG_VU; ++u) {
        const int i = min(i0 + u * PCG_THREADS, P.vhi - 1);
        xv[u] = x[i];
        pv[u] = p[i];
        rv[u] = r[i];
        qv[u] = q[i];
        cv[u] = pc[i];
      }
#pragma unroll
      for (int u = 0; u < PCG_VU; ++u) {
        const int i = i0 + u * PCG_THREADS;
        if (i < P.vhi) {
          x[i] = xv[u] + alpha * pv[u];
          const T ri = rv[u] - alpha * qv[u];
          r[i] = ri;
          nr2 += (double)ri * (double)ri;
          nrho += (double)ri * (double)cv[u] * (double)ri;
        }
      }
    }
    rho2 = rho1;
    const double2 t = psum2(nr2, nrho, red, P);
    r2 = t.x;
    rho1 = t.y;
  }
  pcg_barrier();
  return (it == step_max) ? -it : it;
}

// The iteration once b (B) and x0 (X) are in place: r = b - A x0, then
// semtex's loop in FP64, or FP64 iterative refinement around a T-precision
// PCG on A d = r.  Vectors: X, R, B, P, Q of npts and LB of nel*next
// doubles; the inner T vectors share the P/Q/LB space.  Returns the
// iteration count (summed over inner solves), negative at the limit.
// OPK: the operator of the inner T-precision iteration; OP64: the one of
// the FP64 residuals.
template <typename T, bool REFINE, int OPK = OPK_BLOCK, int OP64 = OPK_BLOCK>
__device__ __forceinline__ int pcg_run(const PcgOp& op, int n, double* X, double* R, const double* B,
                       double* P, double* Q, double* LB, const double* pc, const float* pc32,
                       double epsb2, int step_max, double inner_tol, int max_outer, T* sh,
                       double* red, PcgPart& pt) {
  const int tid = threadIdx.x, npts = op.npts;
  pcg_barrier();
  pcg_op<double, true, OP64>(op, n, X, Q, LB, X, (double*)sh, red, pt);
  for (int i = pt.vlo + tid; i < pt.vhi; i += PCG_THREADS) R[i] = B[i] - Q[i];
  pcg_barrier();
  int its;
  if constexpr (!REFINE) {
    its = pcg_core<double, OP64>(op, n, X, R, P, Q, LB, pc, epsb2, step_max, (double*)sh, red,
                                  pt);
  } else {
    T* d = (T*)P;
    T* rr = d + npts;
    T* pp = rr + npts;
    T* qq = pp + npts;
    T* lb = qq + npts;
    its = 0;
    for (int outer = 0;; ++outer) {
      double rn = 0.0;
      for (int i = pt.vlo + tid; i < pt.vhi; i += PCG_THREADS) rn += R[i] * R[i];
      rn = psum(rn, red, pt);
      if (rn <= epsb2) break;
      if (outer == max_outer) { its = -its; break; }
      for (int i = pt.vlo + tid; i < pt.vhi; i += PCG_THREADS) {
        d[i] = (T)0;
        rr[i] = (T)R[i];
      }
      pcg_barrier();
      const double ein = inner_tol * inner_tol * rn;
      const int it = pcg_core<T, OPK>(op, n, d, rr, pp, qq, lb, (const T*)pc32, ein, step_max,
                                      sh, red, pt);
      its += (it < 0) ? -it : it;
      for (int i = pt.vlo + tid; i < pt.vhi; i += PCG_THREADS) X[i] += (double)d[i];
      pcg_barrier();
      pcg_op<double, true, OP64>(op, n, X, Q, LB, X, (double*)sh, red, pt);
      for (int i = pt.vlo + tid; i < pt.vhi; i += PCG_THREADS) R[i] = B[i] - Q[i];
      pcg_barrier();
    }
  }
  return its;
}

// x0: the current field, essential and pinned entries zeroed
// (Field::local2global, src/field.cpp:928-947).  Where element copies of a
// shared node differ, the copy written last wins (Veclib::gathr_scatr over
// elements in order, src/element.cpp:323): the last entry of the node's
// assembly list, which is in increasing element-storage order.
__device__ __forceinline__ void pcg_x0(const PcgOp& op, const int* emap, int n, const double* uk,
                              double* X, const PcgPart& P) {
  const int tid = threadIdx.x, next = op.next, nint = op.nint;
  for (int gI = P.vlo + tid; gI < min(P.vhi, op.nglobal); gI += PCG_THREADS) {
    const int c = op.asm_idx[op.asm_ptr[gI + 1] - 1];
    const int e = c / next, b = c - (c / next) * next;
    X[gI] = (gI < n) ? uk[(size_t)e * op.npnp + emap[b]] : 0.0;
  }
  for (int i = max(P.vlo - op.nglobal, 0) + tid; i < P.vhi - op.nglobal; i += PCG_THREADS) {
    const int e = i / nint, q = i - (i / nint) * nint;
    X[op.nglobal + i] = uk[(size_t)e * op.npnp + emap[next + q]];
  }
}

// unknowns back to the plane with the essential values re-imposed
// (Field::global2local + setEssential, src/field.cpp:950-969, 1022-1040);
// ess(gid) gives the value of essential node gid >= nsolve.
template <class Ess>
__device__ __forceinline__ void pcg_store(const PcgOp& op, int n, int nsolve, const double* X,
                                 double* uk, Ess ess, const PcgPart& P) {
  const int next = op.next, npnp = op.npnp;
  for (int idx = P.elo * npnp + threadIdx.x; idx < P.ehi * npnp; idx += PCG_THREADS) {
    const int e = idx / npnp, q = idx - (idx / npnp) * npnp;
    const int pm = op.pmap[q];
    double v;
    if (pm < next) {
      const int gid = op.btog[e * next + pm];
      v = (gid < n) ? X[gid] : (gid >= nsolve ? ess(gid) : 0.0);
    } else {
      v = X[op.nglobal + e * op.nint + (pm - next)];
    }
    uk[idx] = v;
  }
}

template <typename T, bool REFINE, int OPK = OPK_BLOCK>
__global__ __launch_bounds__(PCG_THREADS) void k_pcg(PcgArgs a) {
  extern __shared__ double pcg_lds[];
  double* red = pcg_lds;               // RED_DOUBLES doubles
  T* sh = (T*)(pcg_lds + RED_DOUBLES);
  const int tid = threadIdx.x;
  int slot, wg, W;
  if (a.wmap) {
    const int* wm = a.wmap + 4 * blockIdx.x;
    slot = wm[0]; wg = wm[1]; W = wm[2];
  } else {
    W = a.W; slot = blockIdx.x / W; wg = blockIdx.x - slot * W;
  }
  slot = uni(slot);
  const int pidx = a.pcg_plane[slot];
  const int k = a.plane_of[pidx];
  const int m = k >> 1;
  const int n = a.mode_n[m];
  const int npts = a.npts, nglobal = a.nglobal, next = a.next;
  const PcgOp op{a.np, a.npnp, next, a.nint, a.nel, nglobal, npts, a.hcon[m],
                 a.lambda2 > EPSDP, a.G, a.eset, a.pmap, a.btog, a.ops, a.asm_ptr, a.asm_idx,
                 a.nset, a.mq ? a.mq + (size_t)m * a.nset * a.npnp : nullptr};
  double* wk = a.work + (size_t)slot * a.workstride;
  PcgPart PP = part_of(W, wg, npts, a.nel, nglobal, a.pctr + (size_t)slot * 16,
                      a.pslot + (size_t)slot * 4 * PCG_WMAX, a.err);
  double* X = wk;                 // FP64 solution
  double* R = X + npts;           // FP64 residual
  double* B = R + npts;           // FP64 right-hand side
  double* P = B + npts;
  double* Q = P + npts;
  double* LB = Q + npts;          // nel*next element-boundary scratch
  const double* fb = a.Fb + (size_t)pidx * a.nel * next;
  const double* fi = a.Fi + (size_t)pidx * a.nel * a.nint;
  double* uk = a.u + (size_t)k * a.planeSize;
  // b: assembled -(M f + H g) + <h,w>, essential and pinned rows zero
  // (Field::buildRHS, src/field.cpp:859-925)
  double bb = 0.0;
  for (int gI = PP.vlo + tid; gI < min(PP.vhi, nglobal); gI += PCG_THREADS) {
    double v = 0.0;
    if (gI < n)
      for (int c = a.asm_ptr[gI]; c < a.asm_ptr[gI + 1]; ++c) v += fb[a.asm_idx[c]];
    B[gI] = v;
    bb += v * v;
  }
  for (int i = max(PP.vlo - nglobal, 0) + tid; i < PP.vhi - nglobal; i += PCG_THREADS) {
    B[nglobal + i] = fi[i];
    bb += fi[i] * fi[i];
  }
  const double epsb = a.tol * sqrt(psum(bb, red, PP));
  pcg_x0(op, a.emap, n, uk, X, PP);
  constexpr int OP64 = (OPK == OPK_WAVE) ? OPK_WAVE : OPK_BLOCK;
  const int its = pcg_run<T, REFINE, OPK, OP64>(op, n, X, R, B, P, Q, LB, a.pc + (size_t)m * npts,
                                     a.pc32 + (size_t)m * npts, epsb * epsb, a.step_max,
                                     a.inner_tol, a.max_outer, sh, red, PP);
  const double* gp = a.g + (size_t)pidx * a.nzero;
  pcg_store(op, n, a.nsolve, X, uk, [&](int gid) { return gp[gid - a.nsolve]; }, PP);
  if (tid == 0 && wg == 0) a.iters[slot] = its;
}

// --- independent 2D problems, one workgroup each, in one launch ----------
struct PcgInst {
  PcgOp op;
  int n, nsolve, nbound;
  const int* emap;
  const int* ess_src;       // nglobal - nsolve: BC-line index of each essential node
  const int* nat_ptr;       // per element: natural boundary sides
  const int* nat_b;
  const int* bside;
  const double* barea;
  const double* pc;
  const float* pc32;
  double* work;             // 5*npts + nel*next doubles
  int opk;                  // element operator of the inner iteration (OPK_*),
                            // chosen per (np, element count) by the batch tuner
};

struct PcgBatchArgs {
  // workgroup -> (problem, workgroup of the problem, workgroups of the
  // problem, barrier slot); null: one workgroup per problem, in order
  const int* wmap;
  unsigned* pctr;           // per multi-workgroup problem: barrier arrivals
  double* pslot;            // per multi-workgroup problem: 2 x 8 x 2 partial sums
  int* err;
  const PcgInst* inst;
  const double* const* f;   // per instance: forcing plane
  const double* const* bc;  // per instance: BC lines (null without boundaries)
  double* const* u;         // per instance: initial guess in, solution out
  int* iters;
  double tol, inner_tol;
  int step_max, max_outer;
};

// Field::solve JACPCG for one plane from the caller's forcing and BC lines,
// every step inside the workgroup: the right-hand side -(M f) - A g + <h,w>
// (Field::constrain and buildRHS, src/field.cpp:728-768, 859-925, with the
// lifting A g of the essential values g as one unmasked operator
// application), x0, the iteration and the essential re-imposition.
template <typename T, bool REFINE, bool MF>
__global__ __launch_bounds__(PCG_THREADS) void k_pcg_batch(PcgBatchArgs ba) {
  extern __shared__ double pcg_lds[];
  double* red = pcg_lds;
  T* sh = (T*)(pcg_lds + RED_DOUBLES);
  const int tid = threadIdx.x;
  const int* wm = ba.wmap ? ba.wmap + 4 * blockIdx.x : nullptr;
  const int pid = wm ? wm[0] : (int)blockIdx.x;
  const int wg = wm ? wm[1] : 0, W = wm ? wm[2] : 1, sl = wm ? wm[3] : 0;
  const PcgInst& I = ba.inst[pid];
  const PcgOp& op = I.op;
  const int npts = op.npts, nglobal = op.nglobal, next = op.next, nint = op.nint;
  const int npnp = op.npnp, np = op.np, n = I.n, nsolve = I.nsolve;
  const double* f = ba.f[pid];
  const double* bc = ba.bc[pid];
  double* uk = ba.u[pid];
  double* X = I.work;
  double* R = X + npts;
  double* B = R + npts;
  double* P = B + npts;
  double* Q = P + npts;
  double* LB = Q + npts;
  PcgPart PP = part_of(W, wg, npts, op.nel, nglobal, ba.pctr + (size_t)sl * 16,
                       ba.pslot + (size_t)sl * 32, ba.err);
  // lifting vector: essential values on essential nodes, zero elsewhere
  for (int i = PP.vlo + tid; i < PP.vhi; i += PCG_THREADS)
    P[i] = (i >= nsolve && i < nglobal) ? bc[I.ess_src[i - nsolve]] : 0.0;
  pcg_barrier();
  pcg_op<double, false, OPK_WAVE>(op, n, P, Q, LB, P, (double*)sh, red, PP);
  // b = -(M f) - A g + <h,w>, rows of essential and pinned nodes zero
  double bb = 0.0;
  for (int gI = PP.vlo + tid; gI < min(PP.vhi, nglobal); gI += PCG_THREADS) {
    double v = 0.0;
    if (gI < n) {
      for (int c = op.asm_ptr[gI]; c < op.asm_ptr[gI + 1]; ++c) {
        const int e = op.asm_idx[c] / next, b = op.asm_idx[c] - (op.asm_idx[c] / next) * next;
        const int q = I.emap[b];
        v -= op.G[((size_t)op.eset[e] * G_N + G_Q4) * npnp + q] * f[(size_t)e * npnp + q];
        const int nm = np - 1;
        for (int cc = I.nat_ptr[e]; cc < I.nat_ptr[e + 1]; ++cc) {
          const int bs = I.nat_b[cc];
          const int ii = (b - I.bside[bs] * nm + next) % next;
          if (ii <= nm) v += bc[bs * np + ii] * I.barea[bs * np + ii];
        }
      }
      v -= Q[gI];
    }
    B[gI] = v;
    bb += v * v;
  }
  for (int i = max(PP.vlo - nglobal, 0) + tid; i < PP.vhi - nglobal; i += PCG_THREADS) {
    const int e = i / nint, q = I.emap[next + (i - (i / nint) * nint)];
    const double v = -op.G[((size_t)op.eset[e] * G_N + G_Q4) * npnp + q] *
                         f[(size_t)e * npnp + q] - Q[nglobal + i];
    B[nglobal + i] = v;
    bb += v * v;
  }
  const double epsb = ba.tol * sqrt(psum(bb, red, PP));
  pcg_x0(op, I.emap, n, uk, X, PP);
  // the problem's element operator (uniform over the workgroup): the BF16
  // matrix-core one for FP32 inner iterations, or the scalar block-wide or
  // wavefront-level one
  int its = 0;
  bool done = false;
  if constexpr (MF) {
    if (I.opk == OPK_MFMA) {
      its = pcg_run<float, REFINE, OPK_MFMA, OPK_WAVE>(op, n, X, R, B, P, Q, LB, I.pc, I.pc32,
                                                       epsb * epsb, ba.step_max, ba.inner_tol,
                                                       ba.max_outer, sh, red, PP);
      done = true;
    }
  }
  if (done) {
  } else if (!MF && I.opk == OPK_BLOCK)
    its = pcg_run<T, REFINE, OPK_BLOCK, OPK_BLOCK>(op, n, X, R, B, P, Q, LB, I.pc, I.pc32,
                                                   epsb * epsb, ba.step_max, ba.inner_tol,
                                                   ba.max_outer, sh, red, PP);
  else
    its = pcg_run<T, REFINE, OPK_WAVE, OPK_WAVE>(op, n, X, R, B, P, Q, LB, I.pc, I.pc32,
                                                 epsb * epsb, ba.step_max, ba.inner_tol,
                                                 ba.max_outer, sh, red, PP);
  pcg_store(op, n, nsolve, X, uk, [&](int gid) { return bc[I.ess_src[gid - nsolve]]; }, PP);
  if (tid == 0 && wg == 0) ba.iters[pid] = its;
}

}  // namespace

namespace semx {

// --- kernel timing: HIP events recorded around launches on their stream ---
struct ProfRec { const char* name; hipEvent_t a, b; };
struct ProfAcc { const char* name; double ms; long long count; };
static bool g_prof = false;
bool prof_enabled() { return g_prof; }
static std::vector<ProfRec> g_pending;
static std::vector<ProfAcc> g_acc;

void prof_begin(hipStream_t s, hipEvent_t* a) {
  *a = nullptr;
  if (!g_prof) return;
  if (hipEventCreate(a) == hipSuccess) (void)hipEventRecord(*a, s);
}

void prof_end(hipStream_t s, const char* name, hipEvent_t a) {
  if (!g_prof || !a) return;
  hipEvent_t b;
  if (hipEventCreate(&b) != hipSuccess) return;
  (void)hipEventRecord(b, s);
  g_pending.push_back({name, a, b});
}

void prof_flush() {
  for (ProfRec& r : g_pending) {
    (void)hipEventSynchronize(r.b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    bool found = false;
    for (ProfAcc& q : g_acc)
      if (!strcmp(q.name, r.name)) { q.ms += ms; q.count++; found = true; break; }
    if (!found) g_acc.push_back({r.name, (double)ms, 1});
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_pending.clear();
}

int mesh_upload(Mesh& M) {
  if (M.on_device) return SEMX_OK;
  int st;
  HIPCHK(hipGetDevice(&M.device));
  std::vector<double> ops;
  ops.insert(ops.end(), M.ops.DV.begin(), M.ops.DV.end());
  ops.insert(ops.end(), M.ops.DT.begin(), M.ops.DT.end());
  ops.insert(ops.end(), M.ops.SDV.begin(), M.ops.SDV.end());
  ops.insert(ops.end(), M.ops.SDT.begin(), M.ops.SDT.end());
  if ((st = dev_upload(&M.d.G, M.G.data(), M.G.size()))) return st;
  if ((st = dev_upload(&M.d.eset, M.eset.data(), M.eset.size()))) return st;
  if ((st = dev_upload(&M.d.emap, M.ops.emap.data(), M.ops.emap.size()))) return st;
  if ((st = dev_upload(&M.d.pmap, M.ops.pmap.data(), M.ops.pmap.size()))) return st;
  if ((st = dev_upload(&M.d.btog, M.btog.data(), M.btog.size()))) return st;
  if ((st = dev_upload(&M.d.emask, M.emask.data(), M.emask.size()))) return st;
  if ((st = dev_upload(&M.d.ops, ops.data(), ops.size()))) return st;
  if ((st = dev_upload(&M.d.asm_ptr, M.asm_ptr.data(), M.asm_ptr.size()))) return st;
  if ((st = dev_upload(&M.d.asm_idx, M.asm_idx.data(), M.asm_idx.size()))) return st;
  if ((st = dev_upload(&M.d.ess_src, M.ess_src.data(), M.ess_src.size()))) return st;
  if ((st = dev_upload(&M.d.nat_ptr, M.nat_ptr.data(), M.nat_ptr.size()))) return st;
  if ((st = dev_upload(&M.d.nat_b, M.nat_b.data(), M.nat_b.size()))) return st;
  if ((st = dev_upload(&M.d.bside, M.bside.data(), M.bside.size()))) return st;
  if ((st = dev_upload(&M.d.barea, M.barea.data(), M.barea.size()))) return st;
  if ((st = dev_upload(&M.d.imass, M.imass.data(), M.imass.size()))) return st;
  M.on_device = true;
  return SEMX_OK;
}

void mesh_release(Mesh& M) {
  dev_free(M.d.G); dev_free(M.d.eset); dev_free(M.d.emap); dev_free(M.d.pmap);
  dev_free(M.d.btog);
  dev_free(M.d.emask); dev_free(M.d.ops); dev_free(M.d.asm_ptr); dev_free(M.d.asm_idx);
  dev_free(M.d.ess_src); dev_free(M.d.nat_ptr); dev_free(M.d.nat_b); dev_free(M.d.bside);
  dev_free(M.d.barea); dev_free(M.d.imass);
  M.on_device = false;
}

// Lanes of the cooperative band solve: the lead owns the first `lead`
// chunks of every panel, each follower `follow` further chunks.  The band
// is split so that the workgroups of all direct modes fill the device.
// SEMX_COOP_LEAD and SEMX_COOP_FOLLOW override the chunk counts.
int coop_setup(Helm& H) {
  H.coop_w = 0;
  const int nd = (int)H.dmode.size();
  if (nd == 0 || H.npass == 0 || !H.coop_mode) return SEMX_OK;
  // chunks of NB update rows; with extended panels the lead also owns the
  // NB rows of the block inverse
  const int nb = H.nb, nch = H.xcoop ? (H.kd + nb - 1) / nb : H.kdp / nb;
  int cus = 256;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      cus = v;
  }
  const int tasks = COOP_TT * (nb >= 64 ? 1 : 2);      // RPT <= 2 (1 at 64 columns)
  const int ring = nb >= 64 ? 16 : COOP_RING;
  int lead = std::min(nch, std::max(1, 128 / nb));     // ~128 panel rows for the lead
  if (const char* e = getenv("SEMX_COOP_LEAD")) lead = std::max(1, std::min(nch, atoi(e)));
  lead = std::min(lead, std::min(tasks / nb - (H.xcoop ? 1 : 0), ring - 2));
  const int rest = nch - lead;
  int follow = 0, nf = 0;
  if (rest > 0) {
    const int lanes = std::max(1, std::min(COOP_WMAX - 1, cus / nd - 1));
    follow = (rest + lanes - 1) / lanes;
    if (const char* e = getenv("SEMX_COOP_FOLLOW")) follow = std::max(1, atoi(e));
    // at most COOP_WMAX - 1 followers
    follow = std::max(follow, (rest + COOP_WMAX - 2) / (COOP_WMAX - 1));
    follow = std::min(follow, std::min(ring - 2, tasks / nb));
    nf = (rest + follow - 1) / follow;
  }
  if (1 + nf > COOP_WMAX)                     // band too wide for the lanes
    return H.bt ? SEMX_OK : SEMX_E_UNSUPPORTED;
  H.coop_w = 1 + nf;
  H.coop_rpt = (std::max(lead + (H.xcoop ? 1 : 0), follow) * nb > COOP_TT) ? 2 : 1;
  H.coop_clo.assign(H.coop_w + 1, 0);
  H.coop_clo[1] = lead;
  for (int w = 1; w < H.coop_w; ++w) H.coop_clo[w + 1] = std::min(nch, H.coop_clo[w] + follow);
  int nblkmax = 0;
  for (int m : H.dmode) nblkmax = std::max(nblkmax, H.modes[m].nblk);
  H.nblkmax = nblkmax;
  // hand-off slots for up to SEMX_MAX_FIELDS fields of two planes
  H.coop_xstride = (size_t)nd * std::max(H.coop_w - 1, 1) * nblkmax * 2 * SEMX_MAX_FIELDS * nb;
  int st;
  if ((st = dev_upload<double>(&H.d.inbox, nullptr, 2 * H.coop_xstride))) return st;
  if ((st = dev_upload<double>(&H.d.part, nullptr, 2 * H.coop_xstride))) return st;
  if ((st = dev_upload(&H.d.clo, H.coop_clo.data(), H.coop_clo.size()))) return st;
  if ((st = dev_upload<int>(&H.d.err, nullptr, 1))) return st;
  return SEMX_OK;
}

int helm_upload(Helm& H) {
  if (H.on_device) return SEMX_OK;
  Mesh& M = *H.M;
  int st;
  if ((st = mesh_upload(M))) return st;
  const int nm = H.nmodes;
  std::vector<int> mn(nm);
  std::vector<long long> oh(nm), ol(nm);
  std::vector<double> hc(nm);
  for (int m = 0; m < nm; ++m) {
    mn[m] = H.modes[m].n;
    oh[m] = (long long)H.modes[m].off_H;
    ol[m] = (long long)H.modes[m].off_L;
    hc[m] = M.cyl ? 1.0 : H.lambda2 + H.modes[m].betak2;   // cylindrical: factors in mq
  }
  if ((st = dev_upload(&H.d.hcon, hc.data(), hc.size()))) return st;
  if (!H.mq.empty() && (st = dev_upload(&H.d.mq, H.mq.data(), H.mq.size()))) return st;
  if ((st = dev_upload(&H.d.hbi, H.hbi.data(), H.hbi.size()))) return st;
  if ((st = dev_upload(&H.d.hii, H.hii.data(), H.hii.size()))) return st;
  if ((st = dev_upload(&H.d.Pf, H.Pf.data(), H.Pf.size()))) return st;
  if ((st = dev_upload(&H.d.Pb, H.Pb.data(), H.Pb.size()))) return st;
  if ((st = dev_upload(&H.d.dmode, H.dmode.data(), H.dmode.size()))) return st;
  if ((st = dev_upload(&H.d.mode_plane, H.mode_plane.data(), H.mode_plane.size()))) return st;
  if ((st = dev_upload(&H.d.Linv, H.Linv.data(), H.Linv.size()))) return st;
  if ((st = dev_upload(&H.d.xrows, H.xrows.data(), H.xrows.size()))) return st;
  if ((st = dev_upload(&H.d.LinvT, H.LinvT.data(), H.LinvT.size()))) return st;
  if ((st = dev_upload(&H.d.mode_n, mn.data(), mn.size()))) return st;
  if ((st = dev_upload(&H.d.mode_offH, oh.data(), oh.size()))) return st;
  if ((st = dev_upload(&H.d.mode_offL, ol.data(), ol.size()))) return st;
  std::vector<int> mk(nm);
  for (int m = 0; m < nm; ++m) mk[m] = H.modes[m].kind;
  if ((st = dev_upload(&H.d.mode_kind, mk.data(), mk.size()))) return st;
  if (!H.pcg_plane.empty()) {
    std::vector<float> pc32(H.pc.begin(), H.pc.end());
    if ((st = dev_upload(&H.d.pc, H.pc.data(), H.pc.size()))) return st;
    if ((st = dev_upload(&H.d.pc32, pc32.data(), pc32.size()))) return st;
    if ((st = dev_upload(&H.d.pcg_plane, H.pcg_plane.data(), H.pcg_plane.size()))) return st;
  }
  if ((st = dev_upload(&H.d.plane_of, H.plane_of.data(), H.plane_of.size()))) return st;
  if ((st = coop_setup(H))) return st;
  if (H.npart > 1) {
    int nd = 0;
    for (int m = 0; m < nm; ++m) nd += H.modes[m].kind == SEMX_DIRECT;
    if ((st = dev_upload(&H.d.spk, H.spk.data(), H.spk.size()))) return st;
    if ((st = dev_upload(&H.d.spk_off, H.spk_off.data(), H.spk_off.size()))) return st;
    if ((st = dev_upload(&H.d.prow, H.prow.data(), H.prow.size()))) return st;
    // tail/head slots for one field of two planes (rounded up to whole
    // groups of 8 modes)
    const size_t slots = (size_t)((nd + 7) / 8 * 8) * 2 * H.npart * 2 * H.kd;
    if ((st = dev_upload<double>(&H.d.pslot, nullptr, slots))) return st;
    if (!H.d.err && (st = dev_upload<int>(&H.d.err, nullptr, 1))) return st;
    std::vector<double>().swap(H.spk);
  }
  // host band copies are no longer needed
  std::vector<double>().swap(H.Pf);
  std::vector<double>().swap(H.Pb);
  std::vector<double>().swap(H.Linv);
  std::vector<double>().swap(H.LinvT);
  H.on_device = true;
  return SEMX_OK;
}

void helm_release(Helm& H) {
  dev_free(H.d.hbi); dev_free(H.d.hii); dev_free(H.d.Pf); dev_free(H.d.Pb);
  dev_free(H.d.dmode); dev_free(H.d.mode_plane);
  dev_free(H.d.Linv); dev_free(H.d.LinvT); dev_free(H.d.xrows); dev_free(H.d.mode_n); dev_free(H.d.mode_offH);
  dev_free(H.d.mode_offL); dev_free(H.d.pc); dev_free(H.d.Fb); dev_free(H.d.Fi);
  dev_free(H.d.g); dev_free(H.d.sol); dev_free(H.d.plane_of);
  dev_free(H.d.pc32); dev_free(H.d.mode_kind); dev_free(H.d.pcg_plane); dev_free(H.d.work);
  dev_free(H.d.hcon); dev_free(H.d.mq);
  dev_free(H.d.iters);
  dev_free(H.d.inbox); dev_free(H.d.part); dev_free(H.d.clo);
  dev_free(H.d.err);
  dev_free(H.d.spk); dev_free(H.d.spk_off); dev_free(H.d.prow); dev_free(H.d.pslot);
  dev_free(H.d.pctr); dev_free(H.d.pslot2); dev_free(H.d.pwmap);
  H.pcg_ws = 0;
  H.pcg_nwg = 0;
  H.nf_cap = 0;
  H.on_device = false;
}

int helm_scratch(Helm& H, int nf, hipStream_t s) {
  if (nf <= H.nf_cap) return SEMX_OK;
  Mesh& M = *H.M;
  HIPCHK(hipStreamSynchronize(s));
  dev_free(H.d.Fb); dev_free(H.d.Fi); dev_free(H.d.g); dev_free(H.d.sol);
  dev_free(H.d.work); dev_free(H.d.iters);
  H.nf_cap = 0;
  const size_t np_ = (size_t)H.nplanes * nf;
  const int nzero = M.nglobal - M.nsolve;
  int st;
  if ((st = dev_upload<double>(&H.d.Fb, nullptr, np_ * M.nel * M.next))) return st;
  if ((st = dev_upload<double>(&H.d.Fi, nullptr, np_ * M.nel * (M.nint ? M.nint : 1)))) return st;
  if ((st = dev_upload<double>(&H.d.g, nullptr, np_ * (nzero ? nzero : 1)))) return st;
  if ((st = dev_upload<double>(&H.d.sol, nullptr, np_ * H.solstride))) return st;
  if (!H.pcg_plane.empty()) {
    const size_t npcg = H.pcg_plane.size() * nf;
    if ((st = dev_upload<double>(&H.d.work, nullptr, npcg * H.workstride))) return st;
    if ((st = dev_upload<int>(&H.d.iters, nullptr, npcg))) return st;
  }
  H.nf_cap = nf;
  return SEMX_OK;
}

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Workgroups per PCG plane from the planes' iteration counts of the last
// solve (it: npcg per field; the largest over the fields is taken).  An
// iteration of a plane on W workgroups is modelled as taking A + 1/W of its
// one-workgroup time (A: the share that does not shrink with W; 0.03 was
// the fastest of 0.35, 0.2, 0.1 and 0.03 on 32x32 quads np=11 Nz=128, in
// FP64 and both refinements); workgroups are handed out one at a time to the plane with the
// longest modelled time among those below the cap (PCG_WMAX, and 6144
// rows each at least) until the grid would no longer be resident as a
// whole: the planes that the model ranks below the longest one still get
// the workgroups that would otherwise stay idle, which bounds the cost of
// the model's error (the refined solves' inner iteration counts leave out
// their outer sweeps).  The planes' workgroups are laid out plane by
// plane, the most-split planes first.
int pcg_balance(Helm& H, const std::vector<int>& it, int nf, const void* kfn, size_t lds) {
  if (H.pcg_frozen) return SEMX_OK;
  const int npcg = (int)H.pcg_plane.size();
  int cus = 256, dev = 0, per_cu = 0;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, PCG_THREADS, lds) != hipSuccess)
    per_cu = 0;
  const long budget = (long)per_cu * cus;
  const int wcap = std::max(1, std::min(PCG_WMAX, H.npts / 6144));
  if (npcg == 0 || budget <= npcg || wcap == 1) return SEMX_OK;
  std::vector<double> cnt(npcg, 1.0);
  for (int f = 0; f < nf; ++f)
    for (int p = 0; p < npcg; ++p)
      cnt[p] = std::max(cnt[p], (double)std::abs(it[(size_t)f * npcg + p]));
  constexpr double A = 0.03;
  std::vector<int> Wp(npcg, 1);
  long tot = npcg;
  while (tot < budget) {
    int b = -1;
    for (int p = 0; p < npcg; ++p)
      if (Wp[p] < wcap && (b < 0 || cnt[p] * (A + 1.0 / Wp[p]) > cnt[b] * (A + 1.0 / Wp[b])))
        b = p;
    if (b < 0) break;
    ++Wp[b];
    ++tot;
  }
  std::vector<int> ord(npcg);
  for (int p = 0; p < npcg; ++p) ord[p] = p;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return Wp[x] > Wp[y]; });
  std::vector<int> wm;
  wm.reserve((size_t)tot * 4);
  for (int p : ord)
    for (int w = 0; w < Wp[p]; ++w) { wm.push_back(p); wm.push_back(w); wm.push_back(Wp[p]); wm.push_back(0); }
  dev_free(H.d.pwmap);
  H.pcg_nwg = 0;
  if (H.pcg_ws < (size_t)npcg * nf || !H.d.pctr) {
    dev_free(H.d.pctr);
    dev_free(H.d.pslot2);
    int st = dev_upload<unsigned>(&H.d.pctr, nullptr, (size_t)npcg * nf * 16);
    if (!st) st = dev_upload<double>(&H.d.pslot2, nullptr, (size_t)npcg * nf * 4 * PCG_WMAX);
    if (!st && !H.d.err) st = dev_upload<int>(&H.d.err, nullptr, 1);
    if (st) return st;
    H.pcg_ws = (size_t)npcg * nf;
  }
  const int st = dev_upload(&H.d.pwmap, wm.data(), wm.size());
  if (st) return st;
  H.pcg_nwg = (int)tot;
  H.pcg_wtop = std::max(2, *std::max_element(Wp.begin(), Wp.end()));
  return SEMX_OK;
}

int helm_solve_fields(Helm& H, int nf, double* const* fs, const double* const* bcs,
                      double* const* us, double tol_rel, int step_max, int* iters,
                      hipStream_t s) {
  Mesh& M = *H.M;
  if (iters) *iters = 0;
  if (nf < 1 || nf > SEMX_MAX_FIELDS || !fs || !us) return SEMX_E_ARG;
  for (int i = 0; i < nf; ++i) {
    if (!fs[i] || !us[i]) return SEMX_E_ARG;
    if (M.nbound && (!bcs || !bcs[i])) return SEMX_E_ARG;
  }
  if (!H.pcg_plane.empty() && (!(tol_rel >= 0.0) || step_max < 1)) return SEMX_E_ARG;
  // a graph captured from this call keeps the PCG workgroup map's buffers
  // and grid: the map is not rebuilt after that
  if (stream_capturing(s)) H.pcg_frozen = true;
  int st = helm_upload(H);
  if (!st) st = helm_scratch(H, nf, s);
  if (st) return st;
  const int nzero = M.nglobal - M.nsolve;
  const int nline = M.nbound * M.np;
  const size_t pFb = (size_t)H.nplanes * M.nel * M.next;
  const size_t pFi = (size_t)H.nplanes * M.nel * (M.nint ? M.nint : 1);
  const size_t pg = (size_t)H.nplanes * (nzero ? nzero : 1);
  const size_t psol = (size_t)H.nplanes * H.solstride;
  hipEvent_t ev;
  if (nzero > 0) {
    dim3 grid((nzero + 255) / 256, H.nplanes);
    for (int i = 0; i < nf; ++i) {
      prof_begin(s, &ev);
      hipLaunchKernelGGL(k_ess_gather, grid, dim3(256), 0, s, bcs[i], M.d.ess_src,
                         H.d.plane_of, H.d.g + i * pg, nzero, nline);
      prof_end(s, "k_ess_gather", ev);
    }
  }
  ElemArgs ea;
  ea.np = M.np; ea.npnp = M.npnp; ea.next = M.next; ea.nint = M.nint; ea.nel = M.nel;
  ea.planeSize = M.planeSize; ea.nline = nline; ea.nsolve = M.nsolve; ea.nzero = nzero;
  ea.nset = M.nset; ea.lambda2 = H.lambda2; ea.beta = H.beta; ea.base = H.base;
  ea.hcon = H.d.hcon;
  ea.mq = H.d.mq;
  ea.G = M.d.G; ea.eset = M.d.eset; ea.emap = M.d.emap; ea.pmap = M.d.pmap; ea.btog = M.d.btog;
  ea.emask = M.d.emask; ea.ops = M.d.ops; ea.nat_ptr = M.d.nat_ptr; ea.nat_b = M.d.nat_b;
  ea.bside = M.d.bside; ea.barea = M.d.barea; ea.hbi = H.d.hbi; ea.hii = H.d.hii;
  ea.plane_of = H.d.plane_of; ea.mode_n = H.d.mode_n; ea.mode_kind = H.d.mode_kind;
  ea.Fb = H.d.Fb; ea.Fi = H.d.Fi;
  ea.g = H.d.g; ea.sol = H.d.sol; ea.solstride = H.solstride;
  ea.f = nullptr; ea.bc = nullptr; ea.u = nullptr;
  // per-field arguments of the element kernels
  auto field_args = [&](int i) {
    ElemArgs e = ea;
    e.f = fs[i]; e.bc = bcs ? bcs[i] : nullptr; e.u = us[i];
    e.Fb = H.d.Fb + i * pFb; e.Fi = H.d.Fi + i * pFi; e.g = H.d.g + i * pg;
    e.sol = H.d.sol + i * psol;
    return e;
  };
  const int ce = ELT_THREADS / M.npnp;
  dim3 eg((M.nel + ce - 1) / ce, H.nplanes);
  const size_t lds_c = (2 * (size_t)M.npnp + 4 * (size_t)ce * M.npnp) * sizeof(double);
  const size_t lds_b = (size_t)ce * (M.next + M.nint) * sizeof(double);
  if (lds_c > 64 * 1024 || lds_b > 64 * 1024) {
    HIPCHK(hipFuncSetAttribute((const void*)k_condense,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_c));
    HIPCHK(hipFuncSetAttribute((const void*)k_backsub,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
  }
  // staged hbi when every element shares one geometric set; elements per
  // block: enough blocks for two per CU, at most 2 ce
  int cge = (M.nel * H.nmodes + 511) / 512;
  cge = std::max(1, std::min(2 * ce, cge));
  const size_t lds_cs = ((size_t)M.next * M.nint + 2 * (size_t)M.npnp +
                         2 * (size_t)cge * M.npnp + 3 * (size_t)CND_WAVES * M.npnp) *
                        sizeof(double);
  const bool cstaged = M.nset == 1 && lds_cs <= 160 * 1024;
  if (cstaged && lds_cs > 64 * 1024)
    HIPCHK(hipFuncSetAttribute((const void*)k_condense_staged,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cs));
  for (int i = 0; i < nf; ++i) {
    ElemArgs e = field_args(i);
    prof_begin(s, &ev);
    if (cstaged)
      hipLaunchKernelGGL(k_condense_staged, dim3((M.nel + cge - 1) / cge, H.nmodes),
                         dim3(ELT_THREADS), lds_cs, s, e, H.d.mode_plane, cge);
    else
      hipLaunchKernelGGL(k_condense, eg, dim3(ELT_THREADS), lds_c, s, e);
    prof_end(s, "k_condense", ev);
  }
  if (!H.dmode.empty()) {
    for (int i = 0; i < nf; ++i) {
      prof_begin(s, &ev);
      hipLaunchKernelGGL(k_assemble, dim3((M.nglobal + 255) / 256, H.nplanes), dim3(256), 0, s,
                         H.d.Fb + i * pFb, M.d.asm_ptr, M.d.asm_idx, H.d.plane_of, H.d.mode_n,
                         H.d.mode_kind, H.d.sol + i * psol, H.solstride, M.nglobal,
                         M.nel * M.next);
      prof_end(s, "k_assemble", ev);
    }
    BandArgs ba;
    ba.Pf = H.d.Pf; ba.Pb = H.d.Pb; ba.Linv = H.d.Linv; ba.LinvT = H.d.LinvT;
    ba.mode_n = H.d.mode_n; ba.offP = H.d.mode_offH; ba.offL = H.d.mode_offL;
    ba.dmode = H.d.dmode; ba.mode_plane = H.d.mode_plane;
    ba.asm_ptr = M.d.asm_ptr; ba.asm_idx = M.d.asm_idx;
    ba.Fb = H.d.Fb; ba.sol = H.d.sol; ba.solstride = H.solstride; ba.kd = H.kd;
    ba.kdp = H.kdp; ba.nel = M.nel; ba.next = M.next;
    ba.ring = 64;
    while (ba.ring < H.kd + 2 * H.nb) ba.ring *= 2;
    ba.fstride = (long long)psol;
    ba.xrows = H.d.xrows;
    ba.xblk = H.xblk;
    int npl = 1;
    for (int m : H.dmode) if (H.mode_plane[2 * m + 1] >= 0) npl = 2;
    ba.npl = npl;
    // all fields in one sweep of the factor when the windows fit in LDS,
    // else one launch per field
    int nr = npl * nf, launches = 1;
    if ((size_t)nr * (ba.ring + H.nb) * sizeof(double) > 160 * 1024 ||
        (nr != 1 && nr != 2 && nr != 4 && nr != 6)) {
      nr = npl;
      launches = nf;
    }
    const size_t lds = (size_t)nr * (ba.ring + H.nb) * sizeof(double);
    if (lds > 160 * 1024 && H.coop_w == 0) return SEMX_E_UNSUPPORTED;
    const void* kfn = nullptr;
    bool part_done = false;
    unsigned bt = (unsigned)H.bt;
    size_t lds_x = lds;
    unsigned grid = (unsigned)H.dmode.size();
    if (H.band_ext) {
      // Extended panels.  Candidate sweeps:
      //   1  LDS-resident right-hand sides (32-column blocks, two threads per
      //      panel row), one workgroup per mode and field (split)
      //   2  LDS ring window, all fields in one workgroup per mode
      //   3  LDS ring window, one workgroup per mode and field (split)
      //   4, 5  as 1 and 3 with one workgroup per mode and plane
      //   6  as 5 with one panel row per thread (16-column blocks)
      //   7  as 3 with one panel row per thread (16-column blocks)
      //   8  partitioned sweep: one workgroup per mode, plane and partition,
      //      the partitions joined through spikes (one field)
      // A split puts a mode's workgroups on one XCD, so they read the same
      // panels through its L2.  The first call for a field count times the
      // candidates and keeps the fastest (SEMX_AUTOTUNE=0 / 2..8 force 1 / 2..8).
      ba.kdp = H.ldx;
      ba.ndm = (int)H.dmode.size();
      const int xbt = std::min(512, (H.ldx + 63) / 64 * 64);
      const int xnp = (H.ldx + xbt - 1) / xbt;
      const size_t xr_b = ((size_t)2 * H.xblk * sizeof(int) + 7) / 8 * 8;
      auto ext_fn = [&](int r) -> const void* {
        const void* f = nullptr;
#define SEMX_XPICK(NB_, NP_, BT_)                                                \
        if (H.nb == NB_ && xnp == NP_ && xbt == BT_)                             \
          f = (r == 6)   ? (const void*)k_band_ext<NB_, NP_, BT_, 6, 1>          \
              : (r == 4) ? (const void*)k_band_ext<NB_, NP_, BT_, 4, 1>          \
              : (r == 2) ? (const void*)k_band_ext<NB_, NP_, BT_, 2, 1>          \
              : (r == 1) ? (const void*)k_band_ext<NB_, NP_, BT_, 1, 1> : nullptr;
        SEMX_XPICK(32, 1, 64) SEMX_XPICK(32, 1, 128) SEMX_XPICK(32, 1, 192)
        SEMX_XPICK(32, 1, 256) SEMX_XPICK(32, 1, 320) SEMX_XPICK(32, 1, 384)
        SEMX_XPICK(32, 1, 448) SEMX_XPICK(32, 1, 512) SEMX_XPICK(32, 2, 512)
        SEMX_XPICK(16, 2, 512)
#undef SEMX_XPICK
        return f;
      };
      int npadmax = 0;
      for (int m : H.dmode) npadmax = std::max(npadmax, H.modes[m].nblk * H.nb);
      const int rbt = H.nb == XNB ? (2 * H.ldx + 63) / 64 * 64 : 0;
      const void* rfn = nullptr;
#define SEMX_RPICK(NB_, BT_)                                                     \
      if (H.nb == NB_ && rbt == BT_)                                             \
        rfn = (npl == 2) ? (const void*)k_band_res<NB_, BT_, 2, 1>               \
                         : (const void*)k_band_res<NB_, BT_, 1, 1>;
      SEMX_RPICK(32, 128) SEMX_RPICK(32, 192) SEMX_RPICK(32, 256) SEMX_RPICK(32, 320)
      SEMX_RPICK(32, 384) SEMX_RPICK(32, 448) SEMX_RPICK(32, 512) SEMX_RPICK(32, 576)
      SEMX_RPICK(32, 640) SEMX_RPICK(32, 704) SEMX_RPICK(32, 768) SEMX_RPICK(32, 832)
      SEMX_RPICK(32, 896) SEMX_RPICK(32, 960) SEMX_RPICK(32, 1024)
#undef SEMX_RPICK
      struct Cand { const void* fn; size_t lds; unsigned bt; int F; int ring; };
      constexpr int NCAND = 8;
      Cand cand[NCAND + 1] = {};
      const int gpad = (ba.ndm + 7) / 8 * 8;
      {
        const size_t l1 = (size_t)npl * npadmax * sizeof(double) + xr_b;
        if (rfn && l1 <= 160 * 1024) cand[1] = {rfn, l1, (unsigned)rbt, nf, npadmax};
        const size_t l2 = (size_t)npl * nf * (ba.ring + H.nb) * sizeof(double) + xr_b;
        const void* f2 = ext_fn(npl * nf);
        if (f2 && l2 <= 160 * 1024) cand[2] = {f2, l2, (unsigned)xbt, 1, ba.ring};
        const size_t l3 = (size_t)npl * (ba.ring + H.nb) * sizeof(double) + xr_b;
        const void* f3 = nf > 1 ? ext_fn(npl) : nullptr;
        if (f3 && l3 <= 160 * 1024) cand[3] = {f3, l3, (unsigned)xbt, nf, ba.ring};
        // per field with one panel row per thread (16-column blocks, panels <= 960 rows)
        if (nf > 1 && npl == 2 && H.nb == 16 && H.ldx > 896 && H.ldx <= 960 && l3 <= 160 * 1024)
          cand[7] = {(const void*)k_band_ext<16, 1, 960, 2, 1>, l3, 960u, nf, ba.ring};
        if (npl == 2) {
          const void* r1 = nullptr;
#define SEMX_RPICK1(NB_, BT_)                                                    \
          if (H.nb == NB_ && rbt == BT_) r1 = (const void*)k_band_res<NB_, BT_, 1, 1>;
          SEMX_RPICK1(32, 128) SEMX_RPICK1(32, 192) SEMX_RPICK1(32, 256)
          SEMX_RPICK1(32, 320) SEMX_RPICK1(32, 384) SEMX_RPICK1(32, 448)
          SEMX_RPICK1(32, 512) SEMX_RPICK1(32, 576) SEMX_RPICK1(32, 640)
          SEMX_RPICK1(32, 704) SEMX_RPICK1(32, 768) SEMX_RPICK1(32, 832)
          SEMX_RPICK1(32, 896) SEMX_RPICK1(32, 960) SEMX_RPICK1(32, 1024)
#undef SEMX_RPICK1
          const size_t l4 = (size_t)npadmax * sizeof(double) + xr_b;
          if (r1 && l4 <= 160 * 1024) cand[4] = {r1, l4, (unsigned)rbt, 2 * nf, npadmax};
          const size_t l5 = (size_t)(ba.ring + H.nb) * sizeof(double) + xr_b;
          const void* f5 = ext_fn(1);
          if (f5 && l5 <= 160 * 1024) cand[5] = {f5, l5, (unsigned)xbt, 2 * nf, ba.ring};
          // one row per thread: 16-column blocks whose panels fit 960 rows
          if (H.nb == 16 && H.ldx > 896 && H.ldx <= 960 && l5 <= 160 * 1024)
            cand[6] = {(const void*)k_band_ext<16, 1, 960, 1, 1>, l5, 960u, 2 * nf, ba.ring};
        }
        // partitioned sweep, one workgroup per mode, plane and partition (one
        // field), if the whole grid can be resident at once (not when the
        // cooperative sweep is configured: that one then runs instead)
        if (nf == 1 && H.npart > 1 && H.d.spk && H.coop_w == 0) {
          const void* pfn = nullptr;
#define SEMX_PPICK(BT_) if (H.nb == 32 && rbt == BT_) pfn = (const void*)k_band_part<32, BT_>;
          SEMX_PPICK(128) SEMX_PPICK(192) SEMX_PPICK(256) SEMX_PPICK(320) SEMX_PPICK(384)
          SEMX_PPICK(448) SEMX_PPICK(512) SEMX_PPICK(576) SEMX_PPICK(640) SEMX_PPICK(704)
          SEMX_PPICK(768) SEMX_PPICK(832) SEMX_PPICK(896) SEMX_PPICK(960) SEMX_PPICK(1024)
#undef SEMX_PPICK
          const int P = H.npart;
          int maxloc = 0;
          for (int m : H.dmode) {
            const int* pr = &H.prow[(size_t)m * (P + 1)];
            const int nbk = H.modes[m].nblk;
            for (int q = 0; q < P; ++q) {
              const int J1 = (q == P - 1) ? nbk : pr[q + 1] / H.nb;
              maxloc = std::max(maxloc, (J1 - pr[q] / H.nb) * H.nb);
            }
          }
          const size_t l8 = ((size_t)maxloc + (size_t)(H.kd + 1) / 2 * 2 + 2 * (size_t)rbt) *
                                sizeof(double) + 8 + xr_b;
          const unsigned g8 = (unsigned)(gpad * npl * P);
          int per_cu = 0, cus = 0, dev = 0;
          if (pfn && l8 <= 160 * 1024 && hipGetDevice(&dev) == hipSuccess &&
              hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
              hipFuncSetAttribute(pfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l8) ==
                  hipSuccess &&
              hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pfn, rbt, l8) == hipSuccess &&
              (long)per_cu * cus >= (long)g8)
            cand[8] = {pfn, l8, (unsigned)rbt, npl, maxloc};
        }
      }
      // the partitioned sweep's launch: hand-off slots back to the sentinel,
      // then one workgroup per mode, plane and partition
      auto launch_part = [&](const BandArgs& bc, const Cand& cd, bool prof) -> int {
        PartArgs pa;
        pa.b = bc;
        pa.spk = H.d.spk; pa.spk_off = H.d.spk_off; pa.prow = H.d.prow;
        pa.slot = H.d.pslot; pa.err = H.d.err; pa.P = H.npart;
        const size_t used = (size_t)gpad * cd.F * H.npart * 2 * H.kd;
        hipLaunchKernelGGL(k_coop_fill, dim3((unsigned)((used + 255) / 256)), dim3(256), 0, s,
                           H.d.pslot, used);
        void* kargs[] = {&pa};
        if (prof) prof_begin(s, &ev);
        HIPCHK(hipLaunchKernel(cd.fn, dim3((unsigned)(gpad * cd.F * H.npart)), dim3(cd.bt),
                               kargs, cd.lds, s));
        if (prof) prof_end(s, "k_band_solve", ev);
        return SEMX_OK;
      };
      int navail = 0, only = 0;
      for (int c = 1; c <= NCAND; ++c)
        if (cand[c].fn) { ++navail; only = c; }
      if (navail == 0) return SEMX_E_UNSUPPORTED;
      int& pick = H.band_pick[nf];
      if (pick == 0 || !cand[pick].fn) {
        const char* at = getenv("SEMX_AUTOTUNE");
        const int force = at ? atoi(at) : -1;
        if (force >= 0 && force <= NCAND) {
          pick = force == 0 ? 1 : force;
          if (!cand[pick].fn) pick = only;
        } else if (navail == 1) {
          pick = only;
        } else {
          // time every candidate on a zeroed copy of the right-hand sides
          double* tune = nullptr;
          int st2 = dev_upload<double>(&tune, nullptr, (size_t)nf * psol);
          if (st2) return st2;
          hipEvent_t e0, e1;
          HIPCHK(hipEventCreate(&e0));
          HIPCHK(hipEventCreate(&e1));
          float best = 0.f;
          for (int c = 1; c <= NCAND; ++c) {
            if (!cand[c].fn) continue;
            BandArgs bc = ba;
            bc.sol = tune;
            bc.ring = cand[c].ring;
            bc.fsplit = cand[c].F;
            const unsigned gc = cand[c].F > 1 ? (unsigned)(gpad * cand[c].F) : (unsigned)bc.ndm;
            HIPCHK(hipFuncSetAttribute(cand[c].fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)cand[c].lds));
            for (int rep = 0; rep < 3; ++rep) {
              if (rep == 1) HIPCHK(hipEventRecord(e0, s));
              if (c == 8) {
                const int st3 = launch_part(bc, cand[c], false);
                if (st3) return st3;
                continue;
              }
              void* kargs[] = {&bc};
              HIPCHK(hipLaunchKernel(cand[c].fn, dim3(gc), dim3(cand[c].bt), kargs,
                                     cand[c].lds, s));
            }
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (pick == 0 || ms < best) { pick = c; best = ms; }
          }
          HIPCHK(hipEventDestroy(e0));
          HIPCHK(hipEventDestroy(e1));
          dev_free(tune);
        }
      }
      const Cand& cc = cand[pick];
      kfn = cc.fn;
      lds_x = cc.lds;
      bt = cc.bt;
      ba.ring = cc.ring;
      ba.fsplit = cc.F;
      grid = cc.F > 1 ? (unsigned)(gpad * cc.F) : (unsigned)ba.ndm;
      nr = pick == 2 ? npl * nf : (pick >= 4 && pick <= 6) ? 1 : npl;
      launches = 1;
      if (pick == 8) {
        HIPCHK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_x));
        const int st3 = launch_part(ba, cc, true);
        if (st3) return st3;
        part_done = true;
      }
    }
    if (!H.band_ext) {
#define SEMX_PICK(NB_, NP_, BT_, D_)                                        \
    if (H.nb == NB_ && H.npass == NP_ && H.bt == BT_)                         \
      kfn = (nr == 6)   ? (const void*)k_band_solve<NB_, NP_, BT_, 6, D_>   \
            : (nr == 4) ? (const void*)k_band_solve<NB_, NP_, BT_, 4, D_>   \
            : (nr == 2) ? (const void*)k_band_solve<NB_, NP_, BT_, 2, D_>   \
                        : (const void*)k_band_solve<NB_, NP_, BT_, 1, D_>;
    SEMX_PICK(32, 1, 64, 1)
    SEMX_PICK(32, 1, 128, 1)
    SEMX_PICK(32, 1, 256, 1)
    SEMX_PICK(32, 1, 512, 1)
    SEMX_PICK(16, 2, 512, 1)
    SEMX_PICK(8, 4, 512, 1)
#undef SEMX_PICK
    }
    if (!kfn && H.coop_w == 0) return SEMX_E_UNSUPPORTED;
    bool done = part_done;
    if (H.coop_w > 0) {
      // cooperative sweep: all fields' planes of a mode in one launch
      const int nrc = npl * nf;
      const void* cfn = nullptr;
#define SEMX_COOP(NB_, NR_)                                                       \
      if (H.nb == NB_ && nrc == NR_)                                              \
        cfn = H.xcoop ? ((H.coop_rpt == 2) ? (const void*)k_band_xcoop<NB_, NR_, 2>    \
                                           : (const void*)k_band_xcoop<NB_, NR_, 1>)   \
                      : ((H.coop_rpt == 2) ? (const void*)k_band_coop<NB_, NR_, 2>     \
                                           : (const void*)k_band_coop<NB_, NR_, 1>);
      SEMX_COOP(32, 1) SEMX_COOP(32, 2) SEMX_COOP(32, 4) SEMX_COOP(32, 6)
      SEMX_COOP(16, 1) SEMX_COOP(16, 2) SEMX_COOP(16, 4) SEMX_COOP(16, 6)
      SEMX_COOP(8, 1) SEMX_COOP(8, 2) SEMX_COOP(8, 4) SEMX_COOP(8, 6)
#undef SEMX_COOP
      if (cfn) {
        CoopArgs ca;
        ca.Pf = H.d.Pf; ca.Pb = H.d.Pb; ca.Linv = H.d.Linv; ca.LinvT = H.d.LinvT;
        ca.mode_n = H.d.mode_n; ca.offP = H.d.mode_offH; ca.offL = H.d.mode_offL;
        ca.dmode = H.d.dmode; ca.mode_plane = H.d.mode_plane; ca.sol = H.d.sol;
        ca.inbox = H.d.inbox; ca.part = H.d.part; ca.clo = H.d.clo;
        ca.err = H.d.err; ca.solstride = H.solstride; ca.kd = H.kd;
        ca.kdp = H.xcoop ? H.ldx : H.kdp;
        ca.ndirect = (int)H.dmode.size(); ca.W = H.coop_w; ca.nblkmax = H.nblkmax;
        ca.npl = npl; ca.fstride = (long long)psol;
        ca.xstride = (long long)H.coop_xstride;
        // every hand-off slot of this launch back to the sentinel
        const size_t used = (size_t)ca.ndirect * std::max(ca.W - 1, 1) * ca.nblkmax * nrc * H.nb;
        for (double* buf : {H.d.inbox, H.d.part})
          for (int dir = 0; dir < 2; ++dir)
            hipLaunchKernelGGL(k_coop_fill, dim3((unsigned)((used + 255) / 256)), dim3(256), 0,
                               s, buf + dir * H.coop_xstride, used);
        const int groups = (ca.ndirect + 7) / 8;
        void* kargs[] = {&ca};
        prof_begin(s, &ev);
        const hipError_t ce2 = hipLaunchCooperativeKernel(
            cfn, dim3((unsigned)(groups * 8 * H.coop_w)),
            dim3((unsigned)(64 * ((nrc * H.nb + 63) / 64) + 64 + COOP_TT)), kargs, 0, s);
        prof_end(s, "k_band_solve", ev);
        if (ce2 != hipSuccess) {
          fprintf(stderr, "semx: cooperative band solve: %s\n", hipGetErrorString(ce2));
          return SEMX_E_HIP;
        }
        done = true;
      }
    }
    if (!done && !kfn) return SEMX_E_UNSUPPORTED;
    if (!done) HIPCHK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_x));
    for (int l = 0; !done && l < launches; ++l) {
      BandArgs bl = ba;
      bl.sol = H.d.sol + (launches > 1 ? l * psol : 0);
      void* kargs[] = {&bl};
      prof_begin(s, &ev);
      HIPCHK(hipLaunchKernel(kfn, dim3(grid), dim3(bt), kargs, lds_x, s));
      prof_end(s, "k_band_solve", ev);
    }
  }
  if (!H.dmode.empty()) {
    // staged operators when every element shares one geometric set
    // elements per block: enough blocks for two per CU, at most 2 ce
    int ge = (int)((M.nel * H.dmode.size() + 511) / 512);
    ge = std::max(1, std::min(2 * ce, ge));
    const size_t lds_s = ((size_t)M.nint * M.nint + (size_t)M.next * M.nint +
                          2 * (size_t)ge * (M.next + M.nint)) * sizeof(double);
    const bool staged = M.nset == 1 && lds_s <= 160 * 1024;
    if (staged && lds_s > 64 * 1024)
      HIPCHK(hipFuncSetAttribute((const void*)k_backsub_staged,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_s));
    for (int i = 0; i < nf; ++i) {
      ElemArgs e = field_args(i);
      prof_begin(s, &ev);
      if (staged)
        hipLaunchKernelGGL(k_backsub_staged, dim3((M.nel + ge - 1) / ge, H.dmode.size()),
                           dim3(ELT_THREADS), lds_s, s, e, H.d.dmode, H.d.mode_plane, ge);
      else
        hipLaunchKernelGGL(k_backsub, eg, dim3(ELT_THREADS), lds_b, s, e);
      prof_end(s, "k_backsub", ev);
    }
  }
  if (!H.pcg_plane.empty()) {
    PcgArgs pa;
    pa.np = M.np; pa.npnp = M.npnp; pa.next = M.next; pa.nint = M.nint; pa.nel = M.nel;
    pa.nglobal = M.nglobal; pa.nsolve = M.nsolve; pa.nzero = nzero; pa.npts = H.npts;
    pa.planeSize = M.planeSize; pa.nset = M.nset;
    pa.lambda2 = H.lambda2; pa.beta = H.beta; pa.base = H.base; pa.hcon = H.d.hcon;
    pa.mq = H.d.mq;
    pa.G = M.d.G; pa.eset = M.d.eset; pa.emap = M.d.emap; pa.pmap = M.d.pmap;
    pa.btog = M.d.btog; pa.ops = M.d.ops; pa.asm_ptr = M.d.asm_ptr; pa.asm_idx = M.d.asm_idx;
    pa.plane_of = H.d.plane_of; pa.pcg_plane = H.d.pcg_plane; pa.mode_n = H.d.mode_n;
    pa.pc = H.d.pc; pa.pc32 = H.d.pc32; pa.Fb = H.d.Fb; pa.Fi = H.d.Fi; pa.g = H.d.g;
    pa.u = nullptr; pa.work = H.d.work; pa.workstride = H.workstride;
    pa.tol = tol_rel; pa.step_max = step_max; pa.inner_tol = H.inner_tol;
    pa.max_outer = H.max_outer; pa.iters = H.d.iters;
    const int cep = PCG_THREADS / M.npnp;
    // LDS of every element operator (block-wide, wavefront-level, BF16)
    size_t lds = RED_DOUBLES * sizeof(double) +
        (2 * (size_t)M.npnp + 3 * (size_t)cep * M.npnp) * sizeof(double);
    lds = std::max(lds, RED_DOUBLES * sizeof(double) +
                        (2 + 3 * (size_t)PCG_WAVES) * M.npnp * sizeof(double));
    if (H.pcg_prec == SEMX_PCG_BF16_REFINED)
      lds = std::max(lds, RED_DOUBLES * sizeof(double) + (size_t)PCG_WAVES * 2 * MF_TILE * sizeof(float));
    auto pcg_fn = [&](int opk) -> const void* {
      if (H.pcg_prec == SEMX_PCG_FP64)
        return opk == OPK_WAVE ? (const void*)k_pcg<double, false, OPK_WAVE>
                               : (const void*)k_pcg<double, false, OPK_BLOCK>;
      if (H.pcg_prec == SEMX_PCG_BF16_REFINED && opk == OPK_MFMA)
        return (const void*)k_pcg<float, true, OPK_MFMA>;
      return opk == OPK_WAVE ? (const void*)k_pcg<float, true, OPK_WAVE>
                             : (const void*)k_pcg<float, true, OPK_BLOCK>;
    };
    const size_t npcg = H.pcg_plane.size();
    // Workgroups per plane: enough that the planes of one field fill the
    // chip's CUs (each 1024-thread workgroup holds a CU), at most 8, each
    // with at least 8192 vector rows, and only if the whole grid can be
    // resident (the plane's workgroups wait for each other); SEMX_PCG_WG
    // sets the count (1: one workgroup per plane).  Once a solve has read
    // back the planes' iteration counts, a balanced map replaces the
    // uniform count (pcg_balance).
    int W = 1;
    const bool bal = H.pcg_nwg > 0 && H.d.pwmap && !getenv("SEMX_PCG_WG");
    {
      int cus = 256, dev = 0, per_cu = 0;
      if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      W = std::max(1, std::min(8, cus / (int)std::max<size_t>(npcg, 1)));
      W = std::max(1, std::min(W, H.npts / 8192));
      if (const char* e = getenv("SEMX_PCG_WG")) W = std::max(1, std::min(8, atoi(e)));
      if (W > 1) {
        const void* probe = pcg_fn(H.pcg_opk < 0 ? OPK_BLOCK : H.pcg_opk);
        if (lds > 64 * 1024)
          HIPCHK(hipFuncSetAttribute(probe, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, probe, PCG_THREADS, lds) !=
                hipSuccess ||
            (long)per_cu * cus < (long)npcg * W)
          W = 1;
      }
      if ((W > 1 || bal) && (H.pcg_ws < npcg * nf || !H.d.pctr)) {
        dev_free(H.d.pctr);
        dev_free(H.d.pslot2);
        int st2 = dev_upload<unsigned>(&H.d.pctr, nullptr, npcg * nf * 16);
        if (!st2) st2 = dev_upload<double>(&H.d.pslot2, nullptr, npcg * nf * 4 * PCG_WMAX);
        if (!st2 && !H.d.err) st2 = dev_upload<int>(&H.d.err, nullptr, 1);
        if (st2) return st2;
        H.pcg_ws = npcg * nf;
      }
      H.pcg_w = bal ? H.pcg_wtop : W;
    }
    pa.W = W;
    pa.wmap = bal ? H.d.pwmap : nullptr;
    const unsigned grid = bal ? (unsigned)H.pcg_nwg : (unsigned)(npcg * W);
    pa.pctr = H.d.pctr;
    pa.pslot = H.d.pslot2;
    pa.err = H.d.err;
    // barrier counters back to zero before a launch (a fill kernel: the
    // launch may be captured into a graph)
    auto zero_ctr = [&](const PcgArgs& pp) {
      if (pp.W > 1 || pp.wmap)
        hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((npcg * 16 + 255) / 256)), dim3(256), 0,
                           s, pp.pctr, npcg * 16);
    };
    if (H.pcg_opk < 0) {
      // The element operator of the system's PCG planes: the BF16 one for
      // SEMX_PCG_BF16_REFINED; otherwise the block-wide or the wavefront-
      // level scalar operator, whichever runs an iteration faster on this
      // system (timed once: PCG_TUNE_ITERS iterations of every plane of the
      // first field into a scratch copy of its solution), or the one
      // SEMX_PCG_OP=block|wave forces.
      const char* env = getenv("SEMX_PCG_OP");
      if (H.pcg_prec == SEMX_PCG_BF16_REFINED) {
        H.pcg_opk = OPK_MFMA;
      } else if (env) {
        const std::string v(env);
        if (v != "block" && v != "wave") return SEMX_E_ARG;
        H.pcg_opk = v == "block" ? OPK_BLOCK : OPK_WAVE;
      } else {
        double* scr = nullptr;
        const size_t nu = (size_t)H.nz * M.planeSize;
        int st2 = dev_upload<double>(&scr, nullptr, nu);
        if (st2) return st2;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double best = 0.0;
        for (int c : {OPK_BLOCK, OPK_WAVE}) {
          PcgArgs pt = pa;
          pt.u = scr;
          pt.tol = 0.0;
          pt.step_max = PCG_TUNE_ITERS;
          pt.max_outer = 1;
          pt.wmap = nullptr;
          const void* fn = pcg_fn(c);
          if (lds > 64 * 1024)
            HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
          HIPCHK(hipMemcpyAsync(scr, us[0], nu * sizeof(double), hipMemcpyDeviceToDevice, s));
          void* kt[] = {&pt};
          zero_ctr(pt);
          HIPCHK(hipEventRecord(e0, s));
          HIPCHK(hipLaunchKernel(fn, dim3((unsigned)(npcg * W)), dim3(PCG_THREADS), kt, lds, s));
          HIPCHK(hipEventRecord(e1, s));
          HIPCHK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          std::vector<int> it(npcg);
          HIPCHK(hipMemcpy(it.data(), H.d.iters, npcg * sizeof(int), hipMemcpyDeviceToHost));
          int mx = 1;
          for (int v : it) mx = std::max(mx, std::abs(v));
          const double per = ms / mx;
          if (c == OPK_BLOCK || per < best) { best = per; H.pcg_opk = c; }
        }
        HIPCHK(hipEventDestroy(e0));
        HIPCHK(hipEventDestroy(e1));
        dev_free(scr);
      }
    }
    const void* kfn = pcg_fn(H.pcg_opk);
    if (lds > 64 * 1024)
      HIPCHK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    for (int i = 0; i < nf; ++i) {
      PcgArgs pf = pa;
      pf.Fb = H.d.Fb + i * pFb; pf.Fi = H.d.Fi + i * pFi; pf.g = H.d.g + i * pg;
      pf.u = us[i];
      pf.work = H.d.work + i * npcg * H.workstride;
      pf.iters = H.d.iters + i * npcg;
      if (W > 1 || bal) {
        pf.pctr = H.d.pctr + (size_t)i * npcg * 16;
        pf.pslot = H.d.pslot2 + (size_t)i * npcg * 4 * PCG_WMAX;
      }
      void* kargs[] = {&pf};
      zero_ctr(pf);
      prof_begin(s, &ev);
      HIPCHK(hipLaunchKernel(kfn, dim3(grid), dim3(PCG_THREADS), kargs, lds, s));
      prof_end(s, "k_pcg", ev);
    }
    HIPCHK(hipGetLastError());
    if (iters) {
      // the step-limit WARNING of src/field.cpp:713 becomes a status
      std::vector<int> it(npcg * nf);
      HIPCHK(hipMemcpyAsync(it.data(), H.d.iters, it.size() * sizeof(int),
                            hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      int mx = 0;
      bool limit = false;
      for (int v : it) { limit |= v < 0; mx = std::max(mx, v < 0 ? -v : v); }
      *iters = mx;
      if (limit) {
        const int st = helm_check(H, s);
        return st ? st : SEMX_E_MAXITER;
      }
      if (!getenv("SEMX_PCG_WG")) {
        const int st = pcg_balance(H, it, nf, kfn, lds);
        if (st) return st;
      }
    }
  }
  HIPCHK(hipGetLastError());
  if (iters) return helm_check(H, s);
  return SEMX_OK;
}

// The cooperative and partitioned band sweeps set a sticky error word when a bounded
// hand-off wait expires (their result is then wrong).  Every host
// synchronisation point of the solve path reads it here, after the stream
// has drained, and clears it, so a timeout is reported exactly once, by the
// first check after it happened, whatever the caller passed for iters.
int helm_check(Helm& H, hipStream_t s) {
  HIPCHK(hipStreamSynchronize(s));
  if ((H.coop_w <= 0 && H.npart <= 1 && H.pcg_w <= 1) || !H.d.err) return SEMX_OK;
  int e = 0;
  HIPCHK(hipMemcpy(&e, H.d.err, sizeof(int), hipMemcpyDeviceToHost));
  if (!e) return SEMX_OK;
  const int zero = 0;
  HIPCHK(hipMemcpy(H.d.err, &zero, sizeof(int), hipMemcpyHostToDevice));
  fprintf(stderr, "semx: cooperative or partitioned band solve: hand-off wait timed out\n");
  return SEMX_E_TIMEOUT;
}

int helm_solve(Helm& H, double* f, const double* bc, double* u, double tol_rel, int step_max,
               int* iters, hipStream_t s) {
  const double* bcs[1] = {bc};
  return helm_solve_fields(H, 1, &f, bcs, &u, tol_rel, step_max, iters, s);
}

}  // namespace semx

// ===========================================================================
// C ABI
// ===========================================================================


extern "C" {

const char* semx_error_string(int status) {
  switch (status) {
    case SEMX_OK: return "ok";
    case SEMX_E_ARG: return "invalid argument";
    case SEMX_E_NOT_SPD: return "failed to factor Helmholtz matrix (not SPD)";
    case SEMX_E_MAXITER: return "PCG step limit exceeded";
    case SEMX_E_HIP: return "HIP runtime error";
    case SEMX_E_NOMEM: return "out of memory";
    case SEMX_E_JACOBIAN: return "element Jacobian nonpositive";
    case SEMX_E_MESH: return "inconsistent mesh connectivity";
    case SEMX_E_UNSUPPORTED: return "unsupported configuration";
    case SEMX_E_TIMEOUT: return "a bounded device-side wait expired (results invalid)";
    case SEMX_E_SCRATCH: return "request larger than the static scratch of an allocation-free call";
    default: return "unknown status";
  }
}

int semx_device_init(int device) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return SEMX_E_ARG;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipFree(nullptr));
  // the device address of the allocation-free calls' static scratch
  return noalloc_scratch(nullptr);
}

int semx_mesh_create(int np, int nel, const double* vx, const double* vy, const int* nbr,
                     const int* skind, int svv_mn, double svv_eps, double kinvis,
                     semx_mesh* mesh) {
  if (!mesh || !vx || !vy || !nbr || !skind) return SEMX_E_ARG;
  *mesh = nullptr;
  semx_mesh_s* h = new (std::nothrow) semx_mesh_s();
  if (!h) return SEMX_E_NOMEM;
  const int st = mesh_build(h->M, np, nel, vx, vy, nbr, skind, svv_mn, svv_eps, kinvis);
  if (st) { delete h; return st; }
  *mesh = h;
  return SEMX_OK;
}

int semx_mesh_create_numbered(int np, int nel, const double* vx, const double* vy,
                              const int* nbr, const int* skind, int svv_mn, double svv_eps,
                              double kinvis, const int* btog, const int* bmask, int nglobal,
                              int nsolve, semx_mesh* mesh) {
  if (!mesh || !vx || !vy || !nbr || !skind || !btog || !bmask || nglobal < 1 || nsolve < 0)
    return SEMX_E_ARG;
  *mesh = nullptr;
  semx_mesh_s* h = new (std::nothrow) semx_mesh_s();
  if (!h) return SEMX_E_NOMEM;
  const int st = mesh_build(h->M, np, nel, vx, vy, nbr, skind, svv_mn, svv_eps, kinvis, btog,
                            bmask, nglobal, nsolve);
  if (st) { delete h; return st; }
  *mesh = h;
  return SEMX_OK;
}

int semx_mesh_create_enumerated(int np, int nel, const double* vx, const double* vy,
                                const int* nbr, const int* skind, int svv_mn, double svv_eps,
                                double kinvis, int optlevel, semx_mesh* mesh) {
  if (!mesh || !vx || !vy || !nbr || !skind || optlevel < 0 || optlevel > 3) return SEMX_E_ARG;
  *mesh = nullptr;
  semx_mesh_s* h = new (std::nothrow) semx_mesh_s();
  if (!h) return SEMX_E_NOMEM;
  const int st = mesh_build(h->M, np, nel, vx, vy, nbr, skind, svv_mn, svv_eps, kinvis, nullptr,
                            nullptr, 0, 0, optlevel);
  if (st) { delete h; return st; }
  *mesh = h;
  return SEMX_OK;
}

int semx_mesh_set_cylindrical(semx_mesh mesh) {
  if (!mesh) return SEMX_E_ARG;
  return mesh_set_cylindrical(mesh->M);
}

int semx_mesh_destroy(semx_mesh mesh) {
  if (!mesh) return SEMX_E_ARG;
  mesh_release(mesh->M);
  delete mesh;
  return SEMX_OK;
}

int semx_mesh_info(semx_mesh mesh, int* info) {
  if (!mesh || !info) return SEMX_E_ARG;
  const Mesh& M = mesh->M;
  const int v[10] = {M.np, M.nel, M.next, M.nint, M.nglobal, M.nsolve, M.nband,
                     M.nbound, M.nPlane, M.planeSize};
  memcpy(info, v, sizeof(v));
  return SEMX_OK;
}

int semx_mesh_nodes(semx_mesh mesh, double* x, double* y) {
  if (!mesh || !x || !y) return SEMX_E_ARG;
  memcpy(x, mesh->M.x.data(), mesh->M.x.size() * sizeof(double));
  memcpy(y, mesh->M.y.data(), mesh->M.y.size() * sizeof(double));
  return SEMX_OK;
}

int semx_mesh_bsides(semx_mesh mesh, int* belmt, int* bside) {
  if (!mesh || (!belmt && mesh->M.nbound) || (!bside && mesh->M.nbound)) return SEMX_E_ARG;
  if (mesh->M.nbound) {
    memcpy(belmt, mesh->M.belmt.data(), mesh->M.nbound * sizeof(int));
    memcpy(bside, mesh->M.bside.data(), mesh->M.nbound * sizeof(int));
  }
  return SEMX_OK;
}

int semx_mesh_numbering(semx_mesh mesh, int* btog, int* bmask, int* emask) {
  if (!mesh || !btog || !bmask || !emask) return SEMX_E_ARG;
  const Mesh& M = mesh->M;
  memcpy(btog, M.btog.data(), M.btog.size() * sizeof(int));
  memcpy(bmask, M.bmask.data(), M.bmask.size() * sizeof(int));
  memcpy(emask, M.emask.data(), M.emask.size() * sizeof(int));
  return SEMX_OK;
}

int semx_mesh_nsets(semx_mesh mesh, int* nsets) {
  if (!mesh || !nsets) return SEMX_E_ARG;
  *nsets = mesh->M.nset;
  return SEMX_OK;
}

}  // extern "C"

static int helm_create_impl(semx_mesh mesh, double lambda2, double beta, int base_mode,
                            int nmodes, int nz, int method, int svv_mz, double svv_epsz,
                            double kinvis, int nmode_total, semx_helm* helm,
                            char field = 'c') {
  if (!mesh || !helm || nmodes < 1 || nz < 1 || base_mode < 0) return SEMX_E_ARG;
  if (!strchr("cpuvw", field) || !field) return SEMX_E_ARG;
  if (method != SEMX_DIRECT && method != SEMX_JACPCG && method != SEMX_MIXED) return SEMX_E_ARG;
  if (!(nz == 1 && nmodes == 1) && nz != 2 * nmodes) return SEMX_E_ARG;
  *helm = nullptr;
  semx_helm_s* h = new (std::nothrow) semx_helm_s();
  if (!h) return SEMX_E_NOMEM;
  Helm& H = h->H;
  H.M = &mesh->M;
  H.lambda2 = lambda2; H.beta = beta; H.base = base_mode; H.nmodes = nmodes;
  H.nz = nz; H.method = method;
  H.svv_mz = svv_mz; H.svv_epsz = svv_epsz; H.svv_nu = kinvis;
  H.svv_nmode = nmode_total;
  H.field = field;
  const int st = helm_build(H);
  if (st) { delete h; return st; }
  *helm = h;
  return SEMX_OK;
}

extern "C" {

int semx_helm_create(semx_mesh mesh, double lambda2, double beta, int base_mode, int nmodes,
                     int nz, int method, semx_helm* helm) {
  return helm_create_impl(mesh, lambda2, beta, base_mode, nmodes, nz, method, -1, 0.0, 1.0, 0,
                          helm);
}

int semx_helm_create_field(semx_mesh mesh, double lambda2, double beta, int base_mode,
                           int nmodes, int nz, int method, int field, semx_helm* helm) {
  if (field <= 0 || field > 127) return SEMX_E_ARG;
  return helm_create_impl(mesh, lambda2, beta, base_mode, nmodes, nz, method, -1, 0.0, 1.0, 0,
                          helm, (char)field);
}

int semx_helm_create_svv(semx_mesh mesh, double lambda2, double beta, int base_mode,
                         int nmodes, int nz, int method, int svv_mz, double svv_epsz,
                         double kinvis, int nmode_total, semx_helm* helm) {
  if (!(kinvis > 0.0) || nmode_total < base_mode + nmodes) return SEMX_E_ARG;
  return helm_create_impl(mesh, lambda2, beta, base_mode, nmodes, nz, method, svv_mz, svv_epsz,
                          kinvis, nmode_total, helm);
}

int semx_helm_destroy(semx_helm helm) {
  if (!helm) return SEMX_E_ARG;
  helm_release(helm->H);
  delete helm;
  return SEMX_OK;
}

int semx_helm_info(semx_helm helm, int* info) {
  if (!helm || !info) return SEMX_E_ARG;
  const Helm& H = helm->H;
  long long bytes = 0;
  for (const ModeSys& S : H.modes)
    if (S.kind == SEMX_DIRECT) bytes += (long long)S.n * H.ldab * 8;
  info[0] = H.nmodes; info[1] = H.nz; info[2] = H.method;
  info[3] = H.modes.empty() ? 0 : H.modes[0].n;
  info[4] = H.ldab; info[5] = H.nb;
  info[6] = (int)(bytes / 1000);
  info[7] = H.modes.empty() ? 0 : H.modes[0].singular;
  return SEMX_OK;
}

int semx_helm_sweep_info(semx_helm helm, int* info) {
  if (!helm || !info) return SEMX_E_ARG;
  const Helm& H = helm->H;
  info[0] = H.npart;
  for (int f = 1; f <= SEMX_MAX_FIELDS; ++f) info[f] = H.band_pick[f];
  info[4] = H.coop_w;
  info[5] = H.pcg_w;
  return SEMX_OK;
}

int semx_helm_element(semx_helm helm, int m, int elmt, double* hbb, double* hbi, double* hii) {
  if (!helm) return SEMX_E_ARG;
  const Helm& H = helm->H;
  const Mesh& M = *H.M;
  if (m < 0 || m >= H.nmodes || elmt < 0 || elmt >= M.nel) return SEMX_E_ARG;
  const size_t ms = (size_t)m * M.nset + M.eset[elmt];
  const int next = M.next, nint = M.nint;
  if (hbb) memcpy(hbb, &H.hbb[ms * next * next], (size_t)next * next * sizeof(double));
  if (hbi && nint) memcpy(hbi, &H.hbi[ms * next * nint], (size_t)next * nint * sizeof(double));
  if (hii && nint) memcpy(hii, &H.hii[ms * nint * nint], (size_t)nint * nint * sizeof(double));
  return SEMX_OK;
}

int semx_helm_solve(semx_helm helm, double* f, const double* bc, double* u, double tol_rel,
                    int step_max, int* iters, void* stream) {
  if (!helm || !f || !u) return SEMX_E_ARG;
  return helm_solve(helm->H, f, bc, u, tol_rel, step_max, iters, (hipStream_t)stream);
}

int semx_helm_iters(semx_helm helm, int* iters, void* stream) {
  if (!helm || !iters) return SEMX_E_ARG;
  Helm& H = helm->H;
  for (int k = 0; k < H.nz; ++k) iters[k] = 0;
  if (H.pcg_plane.empty() || !H.d.iters) return SEMX_OK;
  hipStream_t s = (hipStream_t)stream;
  std::vector<int> it(H.pcg_plane.size());
  HIPCHK(hipMemcpyAsync(it.data(), H.d.iters, it.size() * sizeof(int), hipMemcpyDeviceToHost,
                        s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t i = 0; i < it.size(); ++i) iters[H.plane_of[H.pcg_plane[i]]] = it[i];
  return SEMX_OK;
}

int semx_helm_check(semx_helm helm, void* stream) {
  if (!helm) return SEMX_E_ARG;
  if (!helm->H.on_device) return SEMX_OK;
  return helm_check(helm->H, (hipStream_t)stream);
}

int semx_helm_solve_fields(semx_helm helm, int nfield, double* const* f, const double* const* bc,
                           double* const* u, double tol_rel, int step_max, int* iters,
                           void* stream) {
  if (!helm) return SEMX_E_ARG;
  return helm_solve_fields(helm->H, nfield, f, bc, u, tol_rel, step_max, iters,
                           (hipStream_t)stream);
}

struct semx_batch_s {
  std::vector<semx_helm> helms;
  std::vector<PcgInst> hinst;      // host copy of inst
  bool tuned = false;              // element operators chosen
  double tune_ms = 0.0;            // host time the choice took
  PcgInst* inst = nullptr;         // device, one per problem
  double* work = nullptr;
  int* wmap = nullptr;             // device: workgroup -> problem, part, parts, slot
  unsigned* pctr = nullptr;        // barrier arrivals of multi-workgroup problems
  double* pslot = nullptr;
  int* err = nullptr;
  int nwg = 0, nmulti = 0;         // workgroups of a launch, multi-workgroup problems
  bool frozen = false;             // captured into a graph: layout kept
  int* iters = nullptr;
  const double** ptrs = nullptr;   // device: f, bc, u pointer arrays
  int count = 0, prec = SEMX_PCG_FP64, max_outer = 12;
  double inner_tol = 1e-5;
  size_t lds = 0;
};

static void batch_free(semx_batch_s* b) {
  if (!b) return;
  dev_free(b->inst); dev_free(b->work); dev_free(b->iters); dev_free(b->ptrs);
  dev_free(b->wmap); dev_free(b->pctr); dev_free(b->pslot); dev_free(b->err);
  delete b;
}

// Launch layout.  A launch lasts as long as its longest problem, so the
// problems whose cost exceeds the per-CU share of the whole batch get
// several workgroups (up to 8, at least 4096 rows each; their workgroups
// wait for each other, so they come first in the grid and together take
// at most one workgroup per CU); the rest follow, one workgroup each,
// largest first.  The cost is expected at creation (unknowns x elements
// per side x np, the iteration count growing with the resolution) and
// measured (unknowns x iterations) once a solve has read back the
// problems' iteration counts.  SEMX_BATCH_WG=1: one workgroup per problem.
static int batch_layout(semx_batch_s* b, const std::vector<double>& cost) {
  const int count = b->count;
  int cus = 256, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  int wmax = 8;
  if (const char* e = getenv("SEMX_BATCH_WG")) wmax = std::max(1, std::min(8, atoi(e)));
  double tot = 0.0, cmax = 0.0;
  for (int i = 0; i < count; ++i) {
    tot += cost[i];
    cmax = std::max(cmax, cost[i]);
  }
  const double target = std::max(tot / cus, cmax / wmax);
  std::vector<int> W(count, 1), ord(count);
  for (int i = 0; i < count; ++i) {
    ord[i] = i;
    const int byrows = std::max(1, b->helms[i]->H.npts / 4096);
    W[i] = std::max(1, std::min({wmax, byrows, (int)std::ceil(cost[i] / target - 1e-9)}));
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  // the multi-workgroup problems must all be resident together
  auto multi_wgs = [&]() {
    long t = 0;
    for (int i = 0; i < count; ++i) if (W[i] > 1) t += W[i];
    return t;
  };
  for (int k = 0; multi_wgs() > cus && k < count; ++k)
    for (int i = 0; i < count && multi_wgs() > cus; ++i)
      if (W[ord[count - 1 - i]] > 1) W[ord[count - 1 - i]] = 1;
  std::vector<int> wm;
  int nm = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int i : ord) {
      if ((W[i] > 1) != (pass == 0)) continue;
      for (int w = 0; w < W[i]; ++w) {
        wm.push_back(i); wm.push_back(w); wm.push_back(W[i]);
        wm.push_back(W[i] > 1 ? nm : 0);
      }
      if (W[i] > 1) ++nm;
    }
  dev_free(b->wmap); dev_free(b->pctr); dev_free(b->pslot);
  b->wmap = nullptr; b->pctr = nullptr; b->pslot = nullptr;
  b->nwg = 0; b->nmulti = 0;
  int st = dev_upload(&b->wmap, wm.data(), wm.size());
  if (!st) st = dev_upload<unsigned>(&b->pctr, nullptr, (size_t)std::max(nm, 1) * 16);
  if (!st) st = dev_upload<double>(&b->pslot, nullptr, (size_t)std::max(nm, 1) * 32);
  if (st) return st;
  b->nwg = (int)(wm.size() / 4);
  b->nmulti = nm;
  return SEMX_OK;
}

int semx_batch_create(int count, const semx_helm* helms, semx_batch* batch) {
  if (count < 1 || !helms || !batch) return SEMX_E_ARG;
  *batch = nullptr;
  semx_batch_s* b = new (std::nothrow) semx_batch_s();
  if (!b) return SEMX_E_NOMEM;
  b->count = count;
  std::vector<PcgInst> inst(count);
  std::vector<size_t> off(count + 1, 0);
  for (int i = 0; i < count; ++i) {
    if (!helms[i]) { batch_free(b); return SEMX_E_ARG; }
    Helm& H = helms[i]->H;
    if (H.nz != 1 || H.nmodes != 1 || H.modes[0].kind != SEMX_JACPCG ||
        H.pcg_prec != helms[0]->H.pcg_prec) {
      batch_free(b);
      return SEMX_E_ARG;
    }
    int st = helm_upload(H);
    if (st) { batch_free(b); return st; }
    off[i + 1] = off[i] + (5 * (size_t)H.npts + (size_t)H.M->nel * H.M->next + 1) / 2 * 2;
    b->helms.push_back(helms[i]);
  }
  const Helm& H0 = helms[0]->H;
  b->prec = H0.pcg_prec; b->inner_tol = H0.inner_tol; b->max_outer = H0.max_outer;
  int st = dev_upload<double>(&b->work, nullptr, off[count]);
  if (!st) st = dev_upload<int>(&b->iters, nullptr, count);
  if (!st) st = dev_upload<const double*>(&b->ptrs, nullptr, 3 * (size_t)count);
  if (st) { batch_free(b); return st; }
  for (int i = 0; i < count; ++i) {
    Helm& H = helms[i]->H;
    Mesh& M = *H.M;
    PcgInst& I = inst[i];
    I.op = PcgOp{M.np, M.npnp, M.next, M.nint, M.nel, M.nglobal, H.npts,
                 M.cyl ? 1.0 : H.lambda2 + H.modes[0].betak2, H.lambda2 > EPSDP, M.d.G,
                 M.d.eset, M.d.pmap, M.d.btog, M.d.ops, M.d.asm_ptr, M.d.asm_idx, M.nset,
                 H.d.mq};
    I.n = H.modes[0].n; I.nsolve = M.nsolve; I.nbound = M.nbound;
    I.emap = M.d.emap; I.ess_src = M.d.ess_src; I.nat_ptr = M.d.nat_ptr; I.nat_b = M.d.nat_b;
    I.bside = M.d.bside; I.barea = M.d.barea; I.pc = H.d.pc; I.pc32 = H.d.pc32;
    I.work = b->work + off[i];
    I.opk = (b->prec == SEMX_PCG_BF16_REFINED) ? OPK_MFMA : OPK_WAVE;
    // wavefront-level element operator: a private slice per wavefront
    if (M.np > 13) { batch_free(b); return SEMX_E_UNSUPPORTED; }
    b->lds = std::max(b->lds, RED_DOUBLES * sizeof(double) +
                      (2 * (size_t)M.npnp + 3 * (size_t)PCG_WAVES * M.npnp) * sizeof(double));
    // the block-wide element operator (a tuner candidate)
    b->lds = std::max(b->lds, RED_DOUBLES * sizeof(double) +
                      (2 + 3 * (size_t)(PCG_THREADS / M.npnp)) * M.npnp * sizeof(double));
    if (b->prec == SEMX_PCG_BF16_REFINED)
      b->lds = std::max(b->lds, RED_DOUBLES * sizeof(double) +
                        (size_t)PCG_WAVES * 2 * MF_TILE * sizeof(float));
  }
  st = dev_upload(&b->inst, inst.data(), inst.size());
  if (st) { batch_free(b); return st; }
  b->hinst = inst;
  std::vector<double> cost(count);
  for (int i = 0; i < count; ++i) {
    const Helm& H = helms[i]->H;
    cost[i] = (double)H.npts * std::sqrt((double)H.M->nel) * H.M->np;
  }
  st = batch_layout(b, cost);
  if (!st && !b->err) st = dev_upload<int>(&b->err, nullptr, 1);
  if (st) { batch_free(b); return st; }
  *batch = b;
  return SEMX_OK;
}

// The batch tuner: for every class of problems with the same np and element
// count, time each element operator that the precision's kernel holds
// (scalar block-wide and wavefront-level; with SEMX_PCG_BF16_REFINED the
// wavefront-level and the BF16 matrix-core one) on the class's first
// problem -- a one-workgroup
// launch of BATCH_TUNE_ITERS iterations from the caller's data into a
// scratch solution -- and keep the fastest per iteration for the class.
// SEMX_BATCH_OP=block|wave|mfma forces one operator for every problem
// (mfma only with BF16 precision, block not with it).  Runs once, on the
// first solve, and synchronises the stream.
constexpr int BATCH_TUNE_ITERS = 40;

static int batch_tune(semx_batch_s& b, const double* const* f, const double* const* bc,
                      double* const* u, const void* kfn, hipStream_t s) {
  b.tuned = true;
  const auto t0 = std::chrono::steady_clock::now();
  // the BF16 batch kernel holds the wavefront-level and the matrix-core
  // operators only (the block-wide one would cost it registers)
  std::vector<int> cands = {OPK_BLOCK, OPK_WAVE};
  if (b.prec == SEMX_PCG_BF16_REFINED) cands = {OPK_WAVE, OPK_MFMA};
  if (const char* e = getenv("SEMX_BATCH_OP")) {
    const std::string v(e);
    const int force = v == "block" ? OPK_BLOCK : v == "wave" ? OPK_WAVE
                      : v == "mfma" ? OPK_MFMA : -1;
    if (force < 0 || (force == OPK_MFMA && b.prec != SEMX_PCG_BF16_REFINED)) return SEMX_E_ARG;
    if (force == OPK_BLOCK && b.prec == SEMX_PCG_BF16_REFINED) return SEMX_E_ARG;
    for (PcgInst& I : b.hinst) I.opk = force;
    HIPCHK(hipMemcpy(b.inst, b.hinst.data(), b.hinst.size() * sizeof(PcgInst),
                     hipMemcpyHostToDevice));
    return SEMX_OK;
  }
  // classes: (np, nel) -> problems
  std::map<std::pair<int, int>, std::vector<int>> cls;
  for (int i = 0; i < b.count; ++i) cls[{b.hinst[i].op.np, b.hinst[i].op.nel}].push_back(i);
  size_t umax = 0;
  for (int i = 0; i < b.count; ++i)
    umax = std::max(umax, (size_t)b.helms[i]->H.M->planeSize);
  PcgInst* dinst = nullptr;
  const double** dptr = nullptr;
  double* du = nullptr;
  int* dit = nullptr;
  int st = dev_upload<PcgInst>(&dinst, nullptr, 1);
  if (!st) st = dev_upload<const double*>(&dptr, nullptr, 3);
  if (!st) st = dev_upload<double>(&du, nullptr, umax);
  if (!st) st = dev_upload<int>(&dit, nullptr, 1);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = SEMX_E_HIP;
  for (auto& kv : cls) {
    if (st) break;
    const int r = kv.second[0];
    const size_t nu = (size_t)b.helms[r]->H.M->planeSize;
    const double* ptr3[3] = {f[r], bc ? bc[r] : nullptr, du};
    if (hipMemcpy(dptr, ptr3, sizeof(ptr3), hipMemcpyHostToDevice) != hipSuccess) {
      st = SEMX_E_HIP;
      break;
    }
    PcgBatchArgs a;
    a.wmap = nullptr; a.pctr = nullptr; a.pslot = nullptr; a.err = nullptr;
    a.inst = dinst; a.f = dptr; a.bc = dptr + 1; a.u = (double* const*)(dptr + 2);
    a.iters = dit; a.tol = 0.0; a.inner_tol = b.inner_tol;
    a.step_max = BATCH_TUNE_ITERS; a.max_outer = 1;
    double best = 0.0;
    int pick = b.hinst[r].opk;
    for (int c : cands) {
      PcgInst I = b.hinst[r];
      I.opk = c;
      int its = 0;
      float ms = 0.f;
      void* kargs[] = {&a};
      if (hipMemcpy(dinst, &I, sizeof(I), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpyAsync(du, u[r], nu * sizeof(double), hipMemcpyDeviceToDevice, s) !=
              hipSuccess ||
          hipEventRecord(e0, s) != hipSuccess ||
          hipLaunchKernel(kfn, dim3(1), dim3(PCG_THREADS), kargs, b.lds, s) != hipSuccess ||
          hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
          hipMemcpy(&its, dit, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
        st = SEMX_E_HIP;
        break;
      }
      const double per = ms / std::max(1, std::abs(its));
      if (c == cands[0] || per < best) { best = per; pick = c; }
    }
    for (int i : kv.second) b.hinst[i].opk = pick;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dev_free(dinst); dev_free(dptr); dev_free(du); dev_free(dit);
  if (st) return st;
  HIPCHK(hipMemcpy(b.inst, b.hinst.data(), b.hinst.size() * sizeof(PcgInst),
                   hipMemcpyHostToDevice));
  b.tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                  .count();
  return SEMX_OK;
}

int semx_batch_destroy(semx_batch batch) {
  if (!batch) return SEMX_E_ARG;
  batch_free(batch);
  return SEMX_OK;
}

int semx_batch_solve(semx_batch batch, const double* const* f, const double* const* bc,
                     double* const* u, double tol_rel, int step_max, int* iters,
                     void* stream) {
  if (!batch || !f || !u || !(tol_rel >= 0.0) || step_max < 1) return SEMX_E_ARG;
  semx_batch_s& b = *batch;
  hipStream_t s = (hipStream_t)stream;
  if (iters) *iters = 0;
  std::vector<const double*> ptrs(3 * (size_t)b.count);
  for (int i = 0; i < b.count; ++i) {
    if (!f[i] || !u[i]) return SEMX_E_ARG;
    if (b.helms[i]->H.M->nbound && (!bc || !bc[i])) return SEMX_E_ARG;
    ptrs[i] = f[i];
    ptrs[b.count + i] = bc ? bc[i] : nullptr;
    ptrs[2 * b.count + i] = u[i];
  }
  if (stream_capturing(s)) b.frozen = true;   // a graph keeps the layout's buffers
  HIPCHK(hipMemcpyAsync(b.ptrs, ptrs.data(), ptrs.size() * sizeof(double*),
                        hipMemcpyHostToDevice, s));
  PcgBatchArgs a;
  a.wmap = b.wmap; a.pctr = b.pctr; a.pslot = b.pslot; a.err = b.err;
  a.inst = b.inst; a.f = b.ptrs; a.bc = b.ptrs + b.count;
  a.u = (double* const*)(b.ptrs + 2 * b.count);
  a.iters = b.iters; a.tol = tol_rel; a.inner_tol = b.inner_tol;
  a.step_max = step_max; a.max_outer = b.max_outer;
  // the matrix-core operator's code (and its registers) only in the BF16
  // batch's kernel
  const void* kfn = (const void*)k_pcg_batch<double, false, false>;
  if (b.prec == SEMX_PCG_FP32_REFINED) kfn = (const void*)k_pcg_batch<float, true, false>;
  if (b.prec == SEMX_PCG_BF16_REFINED) kfn = (const void*)k_pcg_batch<float, true, true>;
  if (b.lds > 64 * 1024)
    HIPCHK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b.lds));
  if (!b.tuned) {
    const int st = batch_tune(b, f, bc, u, kfn, s);
    if (st) return st;
  }
  void* kargs[] = {&a};
  hipEvent_t ev;
  // the multi-workgroup problems' barrier counters back to zero (a fill
  // kernel: the launch may be captured into a graph); they are resident
  // together only if the kernel takes one CU per workgroup or less
  if (b.nmulti > 0) {
    int cus = 256, dev = 0, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, PCG_THREADS, b.lds) !=
            hipSuccess || per_cu < 1)
      return SEMX_E_HIP;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((b.nmulti * 16 + 255) / 256)), dim3(256), 0,
                       s, b.pctr, (size_t)b.nmulti * 16);
  }
  prof_begin(s, &ev);
  HIPCHK(hipLaunchKernel(kfn, dim3((unsigned)b.nwg), dim3(PCG_THREADS), kargs, b.lds, s));
  prof_end(s, "k_pcg_batch", ev);
  HIPCHK(hipGetLastError());
  if (iters) {
    std::vector<int> it(b.count);
    HIPCHK(hipMemcpyAsync(it.data(), b.iters, it.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int mx = 0;
    bool limit = false;
    for (int v : it) { limit |= v < 0; mx = std::max(mx, v < 0 ? -v : v); }
    *iters = mx;
    if (b.nmulti > 0) {   // a multi-workgroup problem's barrier wait timed out
      int e = 0;
      HIPCHK(hipMemcpy(&e, b.err, sizeof(int), hipMemcpyDeviceToHost));
      if (e) {
        const int zero = 0;
        HIPCHK(hipMemcpy(b.err, &zero, sizeof(int), hipMemcpyHostToDevice));
        return SEMX_E_TIMEOUT;
      }
    }
    if (limit) return SEMX_E_MAXITER;
    if (b.frozen) return SEMX_OK;
    std::vector<double> cost(b.count);
    for (int i = 0; i < b.count; ++i)
      cost[i] = (double)b.helms[i]->H.npts * std::max(1, std::abs(it[i]));
    return batch_layout(&b, cost);
  }
  return SEMX_OK;
}

int semx_batch_info(semx_batch batch, int* opk, double* tune_ms) {
  if (!batch) return SEMX_E_ARG;
  if (opk)
    for (int i = 0; i < batch->count; ++i) opk[i] = batch->tuned ? batch->hinst[i].opk : -1;
  if (tune_ms) *tune_ms = batch->tune_ms;
  return SEMX_OK;
}

int semx_helm_set_pcg(semx_helm helm, int precision, double inner_tol, int max_outer) {
  if (!helm) return SEMX_E_ARG;
  if (precision != SEMX_PCG_FP64 && precision != SEMX_PCG_FP32_REFINED &&
      precision != SEMX_PCG_BF16_REFINED)
    return SEMX_E_ARG;
  const bool refine = precision != SEMX_PCG_FP64;
  if (refine && (!(inner_tol > 0.0 && inner_tol < 1.0) || max_outer < 1)) return SEMX_E_ARG;
  Helm& H = helm->H;
  // the matrix-core operator pads elements to 16 x 16
  if (precision == SEMX_PCG_BF16_REFINED && H.M->np > 16) return SEMX_E_UNSUPPORTED;
  if (H.pcg_prec != precision) H.pcg_opk = -1;
  H.pcg_prec = precision;
  if (refine) { H.inner_tol = inner_tol; H.max_outer = max_outer; }
  return SEMX_OK;
}

int semx_malloc(void** ptr, size_t bytes) {
  if (!ptr) return SEMX_E_ARG;
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  return SEMX_OK;
}

int semx_free(void* ptr) {
  HIPCHK(hipFree(ptr));
  return SEMX_OK;
}

int semx_memcpy_htod(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SEMX_OK;
}

int semx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SEMX_OK;
}

int semx_profile_enable(int on) {
  prof_flush();
  g_prof = on != 0;
  return SEMX_OK;
}

int semx_profile_reset(void) {
  prof_flush();
  g_acc.clear();
  return SEMX_OK;
}

int semx_profile_query(const char* kernel, double* total_ms, long long* count) {
  if (!kernel || !total_ms || !count) return SEMX_E_ARG;
  prof_flush();
  *total_ms = 0.0;
  *count = 0;
  for (const ProfAcc& q : g_acc)
    if (!strcmp(q.name, kernel)) { *total_ms = q.ms; *count = q.count; }
  return SEMX_OK;
}

int semx_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return SEMX_OK;
}

}  // extern "C"
