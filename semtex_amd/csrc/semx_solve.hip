// semx_solve.hip -- CDNA4 (gfx950) kernels and the C ABI of libsemx.
//
// Direct (static-condensation) Helmholtz solve over all planes of a field,
// the work of semtex's Field::solve DIRECT branch (src/field.cpp:588-625):
//
//   k_ess_gather   essential BC values of every plane   (Field::getEssential)
//   k_condense     -(M f + H g), interior elimination F_b -= hbi F_i, <h,w>
//                  (Field::constrain, Element::local2globalSumSC, Natural::sum)
//   k_band_solve   assembly of the condensed RHS by owner gather, then the
//                  banded Cholesky triangular solves U^T y = b, U x = y with
//                  the whole RHS resident in LDS             (LAPACK dpbtrs)
//   k_backsub      interior recovery u_i = hii F_i - hbi^T u_b and essential
//                  re-imposition (Element::global2localSC, Field::setEssential)
//
// One workgroup owns one plane in k_band_solve: the factor of a mode is
// streamed from HBM/L2 in blocks of NB columns; each block is a dense
// NB x NB diagonal solve (precomputed triangular inverse) and a
// right-looking update of the next kd rows, so every step is a parallel
// dot product across the 1024 threads of the workgroup.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>

#include "semx_internal.h"

using namespace semx;

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "semx: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return SEMX_E_HIP;                                                 \
    }                                                                    \
  } while (0)

namespace {

constexpr int NB = 32;         // band-solve block size (columns per step)
constexpr int BAND_THREADS = 1024;
constexpr int ELT_THREADS = 256;  // element kernels: np*np <= 256

struct ElemArgs {
  int np, npnp, next, nint, nel, planeSize, nline, nsolve, nzero, nset;
  double lambda2, beta;
  int base;
  const double* G;
  const int* eset;
  const int* emap;
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

__global__ __launch_bounds__(ELT_THREADS) void k_condense(ElemArgs a) {
  const int e = blockIdx.x, p = blockIdx.y, t = threadIdx.x;
  const int k = a.plane_of[p];
  const int np = a.np, npnp = a.npnp, next = a.next, nint = a.nint;
  const int m = k >> 1;
  const double bk = a.beta * (a.base + m);
  const double hcon = a.lambda2 + bk * bk;
  const bool svv = a.lambda2 > EPSDP;
  __shared__ double sD[ELT_THREADS], sT[ELT_THREADS], sU[ELT_THREADS];
  __shared__ double sR[ELT_THREADS], sS[ELT_THREADS], sF[ELT_THREADS];
  const int set = a.eset[e];
  const double* g = a.G + (size_t)set * G_N * npnp;
  if (t < npnp) {
    sD[t] = a.ops[(svv ? 2 : 0) * npnp + t];
    sT[t] = a.ops[(svv ? 3 : 1) * npnp + t];
  }
  double F = 0.0;
  if (t < npnp) F = g[G_Q4 * npnp + t] * a.f[(size_t)k * a.planeSize + (size_t)e * npnp + t];
  if (a.emask[e]) {
    if (t < npnp) sU[t] = 0.0;
    __syncthreads();
    if (t < next) {
      const int gid = a.btog[e * next + t];
      sU[a.emap[t]] = (gid >= a.nsolve) ? a.g[(size_t)p * a.nzero + (gid - a.nsolve)] : 0.0;
    }
    __syncthreads();
    const int i = t / np, j = t - (t / np) * np;
    if (t < npnp) {
      double r = 0.0, s = 0.0;
      for (int q = 0; q < np; ++q) {
        r += sU[i * np + q] * sT[q * np + j];
        s += sD[i * np + q] * sU[q * np + j];
      }
      const double q1 = g[G_Q1 * npnp + t], q2 = g[G_Q2 * npnp + t], q3 = g[G_Q3 * npnp + t];
      sR[t] = q1 * r + q3 * s;
      sS[t] = q2 * s + q3 * r;
    }
    __syncthreads();
    if (t < npnp) {
      double T = g[G_Q4 * npnp + t] * sU[t] * hcon;
      double s1 = 0.0, s2 = 0.0;
      for (int q = 0; q < np; ++q) {
        s1 += sT[i * np + q] * sS[q * np + j];
        s2 += sR[i * np + q] * sD[q * np + j];
      }
      T += s1;
      T += s2;
      F += T;
    }
  }
  if (t < npnp) sF[t] = -F;
  __syncthreads();
  const size_t pe = (size_t)p * a.nel + e;
  if (t < nint) a.Fi[pe * nint + t] = sF[a.emap[next + t]];
  if (t < next) {
    const double* hbi = a.hbi + ((size_t)m * a.nset + set) * next * nint;
    double s = 0.0;
    for (int q = 0; q < nint; ++q) s += hbi[t * nint + q] * sF[a.emap[next + q]];
    double fb = sF[a.emap[t]] - s;
    const int nm = np - 1;
    for (int c = a.nat_ptr[e]; c < a.nat_ptr[e + 1]; ++c) {
      const int b = a.nat_b[c];
      const int side = a.bside[b];
      const int i = (t - side * nm + next) % next;
      if (i <= nm) fb += a.bc[(size_t)k * a.nline + b * np + i] * a.barea[b * np + i];
    }
    a.Fb[pe * next + t] = fb;
  }
}

struct BandArgs {
  const double* Hr;
  const double* Hc;
  const double* Linv;
  const double* LinvT;
  const int* mode_n;
  const long long* offH;
  const long long* offL;
  const int* plane_of;
  const int* asm_ptr;
  const int* asm_idx;
  const double* Fb;
  double* sol;
  int solstride, kd, ldab, nel, next, ylen;
};

__device__ inline double half_wave_sum(double v) {
  // reduction over aligned groups of 32 lanes of a 64-wide wavefront
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 1);
  return v;
}

__global__ __launch_bounds__(BAND_THREADS) void k_band_solve(BandArgs a) {
  extern __shared__ double ylds[];
  const int p = blockIdx.x, tid = threadIdx.x;
  const int k = a.plane_of[p];
  const int m = k >> 1;
  const int n = a.mode_n[m];
  const int kd = a.kd, ldab = a.ldab;
  const int nblk = (n + NB - 1) / NB;
  double* y = ylds + kd;  // y[-kd .. nblk*NB + kd + NB)
  for (int i = tid; i < a.ylen; i += BAND_THREADS) ylds[i] = 0.0;
  __syncthreads();
  // condensed RHS: owner gather of the element-boundary contributions
  const double* Fb = a.Fb + (size_t)p * a.nel * a.next;
  for (int gI = tid; gI < n; gI += BAND_THREADS) {
    double s = 0.0;
    for (int c = a.asm_ptr[gI]; c < a.asm_ptr[gI + 1]; ++c) s += Fb[a.asm_idx[c]];
    y[gI] = s;
  }
  const double* Hr = a.Hr + a.offH[m];
  const double* Hc = a.Hc + a.offH[m];
  const double* L = a.Linv + a.offL[m];
  const double* LT = a.LinvT + a.offL[m];
  const int ra = tid >> 5, rq = tid & 31;  // diagonal-block thread layout
  // -- forward: U^T y = b
  for (int J = 0; J < nblk; ++J) {
    const int j0 = J * NB;
    __syncthreads();
    double v = L[(size_t)J * NB * NB + ra * NB + rq] * y[j0 + rq];
    v = half_wave_sum(v);
    __syncthreads();
    if (rq == 0) y[j0 + ra] = v;
    __syncthreads();
    for (int r = tid; r < kd; r += BAND_THREADS) {
      const int i = j0 + NB + r;
      if (i >= n) break;
      double acc = 0.0;
      const double* h = Hr + (size_t)j0 * ldab + NB + r;
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        const int d = NB + r - jj;
        if (d <= kd) acc += h[(size_t)jj * ldab - jj] * y[j0 + jj];
      }
      y[i] -= acc;
    }
  }
  // -- backward: U x = y
  for (int J = nblk - 1; J >= 0; --J) {
    const int j0 = J * NB;
    __syncthreads();
    double v = LT[(size_t)J * NB * NB + ra * NB + rq] * y[j0 + rq];
    v = half_wave_sum(v);
    __syncthreads();
    if (rq == 0) y[j0 + ra] = v;
    __syncthreads();
    for (int r = tid; r < kd; r += BAND_THREADS) {
      const int i = j0 - 1 - r;
      if (i < 0) break;
      double acc = 0.0;
      const double* h = Hc + (size_t)j0 * ldab + kd - 1 - r;
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        if (jj + 1 + r <= kd) acc += h[(size_t)jj * ldab - jj] * y[j0 + jj];
      }
      y[i] -= acc;
    }
  }
  __syncthreads();
  double* sol = a.sol + (size_t)p * a.solstride;
  for (int i = tid; i < n; i += BAND_THREADS) sol[i] = y[i];
}

__global__ __launch_bounds__(ELT_THREADS) void k_backsub(ElemArgs a) {
  const int e = blockIdx.x, p = blockIdx.y, t = threadIdx.x;
  const int k = a.plane_of[p];
  const int next = a.next, nint = a.nint, npnp = a.npnp;
  const int m = k >> 1;
  const int n = a.mode_n[m];
  __shared__ double sub[ELT_THREADS], sFi[ELT_THREADS];
  const size_t pe = (size_t)p * a.nel + e;
  double* out = a.u + (size_t)k * a.planeSize + (size_t)e * npnp;
  if (t < next) {
    const int gid = a.btog[e * next + t];
    const double ub = (gid < n) ? a.sol[(size_t)p * a.solstride + gid] : 0.0;
    sub[t] = ub;
    out[a.emap[t]] = (gid < n) ? ub
                   : (gid >= a.nsolve ? a.g[(size_t)p * a.nzero + (gid - a.nsolve)] : 0.0);
  }
  if (t < nint) sFi[t] = a.Fi[pe * nint + t];
  __syncthreads();
  if (t < nint) {
    const int set = a.eset[e];
    const double* hii = a.hii + ((size_t)m * a.nset + set) * nint * nint;
    const double* hbi = a.hbi + ((size_t)m * a.nset + set) * next * nint;
    double s1 = 0.0, s2 = 0.0;
    for (int c = 0; c < nint; ++c) s1 += hii[t * nint + c] * sFi[c];
    for (int b = 0; b < next; ++b) s2 += hbi[b * nint + t] * sub[b];
    out[a.emap[next + t]] = s1 - s2;
  }
}

// --- kernel timing: HIP events recorded around launches on their stream ---
struct ProfRec { const char* name; hipEvent_t a, b; };
struct ProfAcc { const char* name; double ms; long long count; };
bool g_prof = false;
std::vector<ProfRec> g_pending;
std::vector<ProfAcc> g_acc;

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

template <class T>
int dev_upload(T** dptr, const T* src, size_t count) {
  *dptr = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void**)dptr, count * sizeof(T)));
  if (src) HIPCHK(hipMemcpy(*dptr, src, count * sizeof(T), hipMemcpyHostToDevice));
  else HIPCHK(hipMemset(*dptr, 0, count * sizeof(T)));
  return SEMX_OK;
}

template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace

namespace semx {

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
  dev_free(M.d.G); dev_free(M.d.eset); dev_free(M.d.emap); dev_free(M.d.btog);
  dev_free(M.d.emask); dev_free(M.d.ops); dev_free(M.d.asm_ptr); dev_free(M.d.asm_idx);
  dev_free(M.d.ess_src); dev_free(M.d.nat_ptr); dev_free(M.d.nat_b); dev_free(M.d.bside);
  dev_free(M.d.barea); dev_free(M.d.imass);
  M.on_device = false;
}

int helm_upload(Helm& H) {
  if (H.on_device) return SEMX_OK;
  Mesh& M = *H.M;
  int st;
  if ((st = mesh_upload(M))) return st;
  const int nm = H.nmodes;
  std::vector<int> mn(nm);
  std::vector<long long> oh(nm), ol(nm);
  for (int m = 0; m < nm; ++m) {
    mn[m] = H.modes[m].n;
    oh[m] = (long long)H.modes[m].off_H;
    ol[m] = (long long)H.modes[m].off_L;
  }
  if ((st = dev_upload(&H.d.hbi, H.hbi.data(), H.hbi.size()))) return st;
  if ((st = dev_upload(&H.d.hii, H.hii.data(), H.hii.size()))) return st;
  if ((st = dev_upload(&H.d.Hr, H.Hr.data(), H.Hr.size()))) return st;
  if ((st = dev_upload(&H.d.Hc, H.Hc.data(), H.Hc.size()))) return st;
  if ((st = dev_upload(&H.d.Linv, H.Linv.data(), H.Linv.size()))) return st;
  if ((st = dev_upload(&H.d.LinvT, H.LinvT.data(), H.LinvT.size()))) return st;
  if ((st = dev_upload(&H.d.mode_n, mn.data(), mn.size()))) return st;
  if ((st = dev_upload(&H.d.mode_offH, oh.data(), oh.size()))) return st;
  if ((st = dev_upload(&H.d.mode_offL, ol.data(), ol.size()))) return st;
  if ((st = dev_upload(&H.d.pc, H.pc.data(), H.pc.size()))) return st;
  if ((st = dev_upload(&H.d.plane_of, H.plane_of.data(), H.plane_of.size()))) return st;
  const size_t np_ = (size_t)H.nplanes;
  const int nzero = M.nglobal - M.nsolve;
  if ((st = dev_upload<double>(&H.d.Fb, nullptr, np_ * M.nel * M.next))) return st;
  if ((st = dev_upload<double>(&H.d.Fi, nullptr, np_ * M.nel * (M.nint ? M.nint : 1)))) return st;
  if ((st = dev_upload<double>(&H.d.g, nullptr, np_ * (nzero ? nzero : 1)))) return st;
  if ((st = dev_upload<double>(&H.d.sol, nullptr, np_ * H.solstride))) return st;
  // host band copies are no longer needed
  std::vector<double>().swap(H.Hr);
  std::vector<double>().swap(H.Hc);
  std::vector<double>().swap(H.Linv);
  std::vector<double>().swap(H.LinvT);
  H.on_device = true;
  return SEMX_OK;
}

void helm_release(Helm& H) {
  dev_free(H.d.hbi); dev_free(H.d.hii); dev_free(H.d.Hr); dev_free(H.d.Hc);
  dev_free(H.d.Linv); dev_free(H.d.LinvT); dev_free(H.d.mode_n); dev_free(H.d.mode_offH);
  dev_free(H.d.mode_offL); dev_free(H.d.pc); dev_free(H.d.Fb); dev_free(H.d.Fi);
  dev_free(H.d.g); dev_free(H.d.sol); dev_free(H.d.plane_of);
  H.on_device = false;
}

}  // namespace semx

// ===========================================================================
// C ABI
// ===========================================================================

struct semx_mesh_s { Mesh M; };
struct semx_helm_s { Helm H; };

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
    default: return "unknown status";
  }
}

int semx_device_init(int device) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return SEMX_E_ARG;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipFree(nullptr));
  return SEMX_OK;
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

int semx_helm_create(semx_mesh mesh, double lambda2, double beta, int base_mode, int nmodes,
                     int nz, int method, semx_helm* helm) {
  if (!mesh || !helm || nmodes < 1 || nz < 1 || base_mode < 0) return SEMX_E_ARG;
  if (method != SEMX_DIRECT && method != SEMX_JACPCG && method != SEMX_MIXED) return SEMX_E_ARG;
  if (!(nz == 1 && nmodes == 1) && nz != 2 * nmodes) return SEMX_E_ARG;
  *helm = nullptr;
  semx_helm_s* h = new (std::nothrow) semx_helm_s();
  if (!h) return SEMX_E_NOMEM;
  Helm& H = h->H;
  H.M = &mesh->M;
  H.lambda2 = lambda2; H.beta = beta; H.base = base_mode; H.nmodes = nmodes;
  H.nz = nz; H.method = method; H.nb = NB;
  const int st = helm_build(H);
  if (st) { delete h; return st; }
  *helm = h;
  return SEMX_OK;
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
  Helm& H = helm->H;
  Mesh& M = *H.M;
  (void)tol_rel; (void)step_max;
  if (iters) *iters = 0;
  for (const ModeSys& S : H.modes)
    if (S.kind != SEMX_DIRECT) return SEMX_E_UNSUPPORTED;
  if (M.nbound && !bc) return SEMX_E_ARG;
  int st = helm_upload(H);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const int nzero = M.nglobal - M.nsolve;
  const int nline = M.nbound * M.np;
  if (nzero > 0) {
    dim3 grid((nzero + 255) / 256, H.nplanes);
    hipEvent_t ev;
    prof_begin(s, &ev);
    hipLaunchKernelGGL(k_ess_gather, grid, dim3(256), 0, s, bc, M.d.ess_src, H.d.plane_of,
                       H.d.g, nzero, nline);
    prof_end(s, "k_ess_gather", ev);
  }
  ElemArgs ea;
  ea.np = M.np; ea.npnp = M.npnp; ea.next = M.next; ea.nint = M.nint; ea.nel = M.nel;
  ea.planeSize = M.planeSize; ea.nline = nline; ea.nsolve = M.nsolve; ea.nzero = nzero;
  ea.nset = M.nset; ea.lambda2 = H.lambda2; ea.beta = H.beta; ea.base = H.base;
  ea.G = M.d.G; ea.eset = M.d.eset; ea.emap = M.d.emap; ea.btog = M.d.btog;
  ea.emask = M.d.emask; ea.ops = M.d.ops; ea.nat_ptr = M.d.nat_ptr; ea.nat_b = M.d.nat_b;
  ea.bside = M.d.bside; ea.barea = M.d.barea; ea.hbi = H.d.hbi; ea.hii = H.d.hii;
  ea.plane_of = H.d.plane_of; ea.mode_n = H.d.mode_n; ea.Fb = H.d.Fb; ea.Fi = H.d.Fi;
  ea.g = H.d.g; ea.sol = H.d.sol; ea.solstride = H.solstride; ea.f = f; ea.bc = bc; ea.u = u;
  dim3 eg(M.nel, H.nplanes);
  hipEvent_t ev;
  prof_begin(s, &ev);
  hipLaunchKernelGGL(k_condense, eg, dim3(ELT_THREADS), 0, s, ea);
  prof_end(s, "k_condense", ev);
  BandArgs ba;
  ba.Hr = H.d.Hr; ba.Hc = H.d.Hc; ba.Linv = H.d.Linv; ba.LinvT = H.d.LinvT;
  ba.mode_n = H.d.mode_n; ba.offH = H.d.mode_offH; ba.offL = H.d.mode_offL;
  ba.plane_of = H.d.plane_of; ba.asm_ptr = M.d.asm_ptr; ba.asm_idx = M.d.asm_idx;
  ba.Fb = H.d.Fb; ba.sol = H.d.sol; ba.solstride = H.solstride; ba.kd = H.kd;
  ba.ldab = H.ldab; ba.nel = M.nel; ba.next = M.next;
  int nblk_max = 0;
  for (const ModeSys& S : H.modes) nblk_max = S.nblk > nblk_max ? S.nblk : nblk_max;
  ba.ylen = H.kd + nblk_max * NB + H.kd + NB;
  const size_t lds = (size_t)ba.ylen * sizeof(double);
  if (lds > 160 * 1024) return SEMX_E_UNSUPPORTED;
  HIPCHK(hipFuncSetAttribute((const void*)k_band_solve,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  prof_begin(s, &ev);
  hipLaunchKernelGGL(k_band_solve, dim3(H.nplanes), dim3(BAND_THREADS), lds, s, ba);
  prof_end(s, "k_band_solve", ev);
  prof_begin(s, &ev);
  hipLaunchKernelGGL(k_backsub, eg, dim3(ELT_THREADS), 0, s, ea);
  prof_end(s, "k_backsub", ev);
  HIPCHK(hipGetLastError());
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
