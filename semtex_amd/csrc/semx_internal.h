// semx_internal.h -- host-side data structures shared by the setup code
// (semx_setup.cpp) and the HIP solve code (semx_solve.hip).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "semx.h"

namespace semx {

constexpr int SEMX_MAX_FIELDS = 3;   // fields per batched solve (u, v, w)
constexpr double EPSDP = 6.0e-14;  // roundoff threshold of veclib/cveclib.h:58
constexpr double EPSSP = 6.0e-7;   // singular-system threshold, cveclib.h:54

// Geometric factor slots of one element set (Element::mapping outputs).
// Geometric factors per set.  Cylindrical meshes: Q1-Q3 carry the radius
// y (Helmholtz equations are premultiplied by it), Q4 does not (it is the
// mass of Element::weight, semtex's Q8), G_Y holds y (zero when Cartesian).
enum { G_Q1 = 0, G_Q2, G_Q3, G_Q4, G_DRDX, G_DSDX, G_DRDY, G_DSDY, G_Y, G_N };

struct Ops {
  int np = 0;
  std::vector<double> z, w, DV, DT, SDV, SDT;  // np, np, np*np x4
  std::vector<int> emap, pmap;                 // np*np
};

// Device-resident copy of a mesh.
struct MeshDev {
  double* G = nullptr;       // nset*G_N*npnp
  int* eset = nullptr;       // nel
  int* emap = nullptr;       // npnp
  int* pmap = nullptr;       // npnp, inverse of emap
  int* btog = nullptr;       // nel*next
  int* emask = nullptr;      // nel
  double* ops = nullptr;     // DV DT SDV SDT, 4*npnp
  int* asm_ptr = nullptr;    // nglobal+1
  int* asm_idx = nullptr;    // nel*next
  int* ess_src = nullptr;    // nglobal-nsolve
  int* nat_ptr = nullptr;    // nel+1
  int* nat_b = nullptr;      // natural boundary indices
  int* bside = nullptr;      // nbound
  double* barea = nullptr;   // nbound*np
  double* imass = nullptr;   // nglobal
};

struct Mesh {
  int np = 0, nel = 0, next = 0, nint = 0, npnp = 0, nPlane = 0, planeSize = 0;
  Ops ops;
  std::vector<double> x, y;          // nel*npnp node coordinates
  std::vector<double> vx, vy;        // nel*4 vertex coordinates
  bool cyl = false;                  // cylindrical coordinates (x axial, y radial)
  bool used = false;                 // a system has been built on the mesh
  int nset = 0;
  std::vector<int> eset;             // element -> geometric-factor set
  std::vector<double> G;             // nset*G_N*npnp
  std::vector<int> setQ3;            // set has nonzero Q3
  int nglobal = 0, nsolve = 0, nband = 0;
  std::vector<int> btog, bmask, emask;
  int nbound = 0;
  std::vector<int> belmt, bside, bkind;
  std::vector<double> barea;         // nbound*np side quadrature weights
  std::vector<double> imass;         // nglobal inverse global mass
  std::vector<int> ess_src;          // essential node -> bc-line index
  std::vector<int> asm_ptr, asm_idx; // global node -> element-boundary slots
  std::vector<int> nat_ptr, nat_b;   // element -> natural boundaries
  bool on_device = false;
  int device = -1;
  MeshDev d;
};

// One block-triangular band factor on the device, per mode.
struct ModeSys {
  int kind = SEMX_DIRECT;     // DIRECT or JACPCG
  double betak2 = 0.0;
  int n = 0;                  // solved unknowns (nsolve, or nsolve-1 pinned)
  int singular = 0;
  size_t off_H = 0;           // offset (doubles) into band arrays
  size_t off_L = 0;           // offset (doubles) into block-inverse arrays
  int nblk = 0;
};

struct HelmDev {
  double* hbi = nullptr;      // nmodes*nset*next*nint
  double* hii = nullptr;      // nmodes*nset*nint*nint
  double* Pf = nullptr;       // forward-sweep panels, all modes
  double* Pb = nullptr;       // backward-sweep panels, all modes
  double* Linv = nullptr;     // (U_JJ^T)^-1 blocks, row-major
  double* LinvT = nullptr;    // transposes
  int* mode_n = nullptr;      // per mode n
  int* dmode = nullptr;       // direct modes, one band-solve workgroup each
  int* mode_plane = nullptr;  // per mode: up to 2 plane-list indices (-1 none)
  long long* mode_offH = nullptr;
  long long* mode_offL = nullptr;
  int* xrows = nullptr;       // extended-panel live rows, see Helm::xrows
  double* pc = nullptr;       // JACPCG inverse diagonals, nmodes*npts
  float* pc32 = nullptr;      // the same in FP32
  int* mode_kind = nullptr;   // per mode: DIRECT or JACPCG
  int* pcg_plane = nullptr;   // plane-list indices solved by PCG
  double* work = nullptr;     // PCG vectors, one slice per PCG plane
  int* iters = nullptr;       // PCG iterations per PCG plane
  // cooperative band solve: hand-off slots, chunk ranges
  double* inbox = nullptr;    // 2 directions x coop_xstride, lead -> followers
  double* part = nullptr;     // 2 directions x coop_xstride, followers -> lead
  int* clo = nullptr;         // W+1
  int* err = nullptr;         // 1: a hand-off wait timed out
  // scratch
  double* Fb = nullptr;       // nplanes*nel*next
  double* Fi = nullptr;       // nplanes*nel*nint
  double* g = nullptr;        // nz*nzero essential values
  double* sol = nullptr;      // nplanes*solstride
  int* plane_of = nullptr;    // nplanes -> field plane
  double* hcon = nullptr;     // per mode: lambda2 + betak2 (betak2 with its SVV weight);
                              // 1 on cylindrical meshes, whose mass factors are in mq
  double* mq = nullptr;       // cylindrical: nmodes*nset*npnp operator mass factors
  // partitioned band sweep: spikes, their per-mode offsets, partition row
  // starts (npart+1 per mode) and the tail/head hand-off slots
  double* spk = nullptr;
  long long* spk_off = nullptr;
  int* prow = nullptr;
  double* pslot = nullptr;
  // PCG planes with several workgroups each: barrier counters (16 per
  // plane, one used) and partial-sum slots (2 x 16 x 2 per plane)
  unsigned* pctr = nullptr;
  double* pslot2 = nullptr;
  // PCG workgroup map (4 ints per workgroup: plane slot, part, parts, 0),
  // balanced by the planes' last iteration counts
  int* pwmap = nullptr;
};

struct Helm {
  Mesh* M = nullptr;
  double lambda2 = 0.0, beta = 0.0;
  int base = 0, nmodes = 0, nz = 0, method = SEMX_DIRECT;
  // field name for Field::modeConstant: on cylindrical meshes the coupled
  // radial and azimuthal fields 'v' / 'w' take beta k + 1 / beta k - 1
  // (1 for mode 0); every other case beta k
  char field = 'c';
  // Fourier-direction SVV (SVV_MZ, SVV_EPSZ, KINVIS; src/svv.cpp:82-123):
  // betak2 of mode k > svv_mz weighted by 1 + epsz/nu exp(-((N-k)/(mz-k))^2)
  // when lambda2 > EPSDP (src/matrix.cpp:78-83); svv_mz < 0: off
  int svv_mz = -1, svv_nmode = 0;   // svv_nmode: Fourier modes of the whole field
  double svv_epsz = 0.0, svv_nu = 1.0;
  int nb = 32;                // band-solve block size (columns per step)
  int npass = 1, bt = 1024;   // band-solve row passes, threads per block
  int band_d = 1;             // band-solve panel prefetch depth (steps); depth 2
                              // measured no faster on the headline workload
  // extended panels: diagonal-block inverses folded into the panels (one
  // product per block step); ldx rows per panel column
  bool band_ext = false;
  int ldx = 0;
  bool xcoop = false;         // cooperative sweep over extended panels
  // band-sweep variant per field count, chosen by timing on first use:
  // 0 undecided, 1 right-hand sides resident in LDS (one workgroup per mode
  // and field), 2 LDS ring window (all fields in one workgroup per mode),
  // 3 ring window split per field, 4 / 5 as 1 / 3 split per plane
  int band_pick[SEMX_MAX_FIELDS + 1] = {0};
  // live rows (32 + envelope height) of each extended panel, forward and
  // backward: xrows[(mode*xblk + J)*2 + dir], dir 0 forward
  std::vector<int> xrows;
  int xblk = 0;
  int parts = 1;              // lanes sharing one update row
  int kd = 0, ldab = 0, kdp = 0;
  std::vector<ModeSys> modes;
  std::vector<double> hbb, hbi, hii;  // host, nmodes*nset*...
  std::vector<double> Pf, Pb, Linv, LinvT;  // host, released after upload
  std::vector<double> pc;                   // nmodes*npts
  // cylindrical meshes: the operator's mass factor per mode, set and node,
  // Q4 y (lambda2 + betak2 / y^2), zero on the axis (src/element.cpp:
  // 1564-1599); empty when Cartesian
  std::vector<double> mq;
  size_t band_doubles = 0, inv_doubles = 0;
  int nplanes = 0, solstride = 0, npts = 0;
  std::vector<int> plane_of;
  std::vector<int> dmode, mode_plane;
  std::vector<int> pcg_plane;         // plane-list indices of JACPCG modes
  int pcg_prec = 0;                   // SEMX_PCG_FP64, _FP32_REFINED or _BF16_REFINED
  int pcg_w = 1;                      // workgroups per PCG plane (last solve)
  size_t pcg_ws = 0;                  // plane-workgroup slots allocated
  int pcg_nwg = 0;                    // workgroups of the balanced map (0: none)
  int pcg_wtop = 1;                   // most workgroups of a plane in that map
  bool pcg_frozen = false;            // a solve was captured: map no longer rebuilt
  int pcg_opk = -1;                   // element operator of the PCG planes (block-wide
                                      // or wavefront-level scalar, BF16 matrix cores),
                                      // -1 until chosen on the first solve
  double inner_tol = 1e-5;            // inner relative tolerance (refined)
  int max_outer = 12;                 // refinement sweeps (refined)
  size_t workstride = 0;              // doubles per PCG plane
  int nf_cap = 0;                     // fields the device scratch holds
  bool coop_mode = false;             // cooperative band solve chosen at setup
  int coop_w = 0;                     // workgroups per mode (0: one-workgroup sweep)
  int coop_rpt = 1;                   // panel rows per task thread
  int nblkmax = 0;
  size_t coop_xstride = 0;            // hand-off slots per direction
  std::vector<int> coop_clo;          // chunk ranges per lane
  // partitioned band sweep (32-column extended panels): each mode's band
  // cut into npart row ranges at block boundaries, swept concurrently and
  // joined through spikes computed at setup; npart = 1: not available.
  // Per mode: prow (npart+1 row starts), spk_off into spk, which holds the
  // forward spikes V_p = L_p^-1 B_p of partitions 1..npart-1 and then the
  // backward spikes W_p = U_p^-1 C_p of partitions 0..npart-2, each n_p x
  // kd column-major (L = U^T the factor, B_p / C_p its coupling blocks)
  int npart = 1;
  std::vector<int> prow;
  std::vector<long long> spk_off;
  std::vector<double> spk;            // host, released after upload
  bool on_device = false;
  HelmDev d;
};

// setup (semx_setup.cpp)
// btog/bmask (nel*next each, or null for the library's own numbering),
// nglobal, nsolve: a caller-supplied numbering (semtex's session.num);
// enum_level 0-3: without one, number as `enumerate -O<enum_level>` does
int mesh_build(Mesh& M, int np, int nel, const double* vx, const double* vy,
               const int* nbr, const int* skind, int svv_mn, double svv_eps,
               double kinvis, const int* btog = nullptr, const int* bmask = nullptr,
               int nglobal = 0, int nsolve = 0, int enum_level = -1);
int helm_build(Helm& H);
// switch a host mesh (not yet on the device) to cylindrical coordinates:
// factors and side weights multiplied by the radius y
int mesh_set_cylindrical(Mesh& M);
void gll(int np, std::vector<double>& z, std::vector<double>& w);
void gll_deriv(int np, const std::vector<double>& z, std::vector<double>& D);
void svv_derivative(int np, int svv_mn, double svv_eps, double kinvis,
                    std::vector<double>& SDV);

// device (semx_solve.hip)
int mesh_upload(Mesh& M);
void mesh_release(Mesh& M);
int helm_upload(Helm& H);
void helm_release(Helm& H);

}  // namespace semx

// the C ABI's opaque handles
struct semx_mesh_s { semx::Mesh M; };
struct semx_helm_s { semx::Helm H; };
