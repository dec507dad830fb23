/*
 * semx.h -- C ABI of the MI355X-native spectral-element / Fourier
 * Helmholtz core (libsemx.so).
 *
 * The ABI replaces the lower boundary of semtex's Helmholtz hot path: the
 * work that semtex's C++ classes (Field, MatrixSys, ModalMatrixSys,
 * AuxField) hand to BLAS/LAPACK, veclib gather/scatter and femlib.  Each
 * entry point names the reference interface it stands in for.  Plain
 * pointers and sizes only; device pointers are HIP device addresses,
 * streams are hipStream_t passed as void* (NULL = default stream).
 *
 * Storage conventions are semtex's:
 *   - element data row-major np x np, e*np*np + i*np + j (i = s row,
 *     j = r column)                                  (src/element.h:23-24)
 *   - a field is nz planes of planeSize doubles, planeSize = nel*np*np
 *     rounded up to even                         (src/geometry.cpp:105-108)
 *   - plane 2m = Re(mode m), 2m+1 = Im(mode m), plane 1 = Nyquist, which
 *     Field::solve skips                  (src/auxfield.h:18-29, field.cpp:566)
 *   - boundary-condition lines: nbound*np values per plane, one group of np
 *     per boundary side in the mesh's boundary list, CCW along the side
 *                                                    (src/field.cpp:77-150)
 *
 * Errors: every call returns a status (SEMX_OK = 0).  semtex prints and
 * exits instead (veclib/util.c:23-45); semx_error_string() gives the text.
 */
#ifndef SEMX_H
#define SEMX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- status codes -------------------------------------------------------- */
#define SEMX_OK           0
#define SEMX_E_ARG        1   /* bad argument / size                       */
#define SEMX_E_NOT_SPD    2   /* factorisation failed: matrix.cpp:261       */
#define SEMX_E_MAXITER    3   /* PCG step limit: field.cpp:706 (WARNING)    */
#define SEMX_E_HIP        4   /* HIP runtime error / no device              */
#define SEMX_E_NOMEM      5   /* allocation failed                          */
#define SEMX_E_JACOBIAN   6   /* nonpositive Jacobian: element.cpp:1358     */
#define SEMX_E_MESH       7   /* inconsistent mesh connectivity             */
#define SEMX_E_UNSUPPORTED 8  /* configuration not supported                */
#define SEMX_E_TIMEOUT    9   /* a bounded device-side wait expired: the    *
                               * results of the solves since the last check *
                               * are invalid (no counterpart in semtex)     */
#define SEMX_E_SCRATCH   10   /* request exceeds the static scratch of an   *
                               * allocation-free call                       */

/* --- solver kinds (femlib/cfemdef.h:43-52 SolverKind) -------------------- */
#define SEMX_DIRECT  0        /* static condensation + banded Cholesky      */
#define SEMX_JACPCG  1        /* Jacobi-preconditioned CG                   */
#define SEMX_MIXED   2        /* mode 0 DIRECT, other modes JACPCG          */

/* --- JACPCG iteration precision (semx_helm_set_pcg) ------------------- */
#define SEMX_PCG_FP64          0   /* FP64 vectors, semtex's loop exactly    */
#define SEMX_PCG_FP32_REFINED  1   /* FP32 inner PCG, FP64 refinement        */
#define SEMX_PCG_BF16_REFINED  2   /* FP32 inner PCG whose element operator   *
                                    * runs on the BF16 matrix cores (np <= 16,*
                                    * padded to 16, FP32 accumulation), FP64  *
                                    * refinement                              */

/* --- transform direction (femlib/cfemdef.h TransformKind) --------------- */
#define SEMX_FORWARD  1
#define SEMX_INVERSE -1

typedef struct semx_mesh_s* semx_mesh;
typedef struct semx_helm_s* semx_helm;
typedef struct semx_dns_s*  semx_dns;
typedef struct semx_batch_s* semx_batch;

/* Text for a status code.                                                  */
const char* semx_error_string (int status);

/* Select the HIP device for subsequent calls of this thread and check that
 * it is usable.                                                            */
int semx_device_init (int device);

/* --- mesh: Element construction + numbering ------------------------------
 * Replaces Element::Element / Element::mapping (src/element.cpp:30-106,
 * 1281-1436), Mesh::meshElmt (src/mesh.cpp:772-886) and the numbering that
 * semtex obtains by running `enumerate` (src/bcmgr.cpp:586-824,
 * utility/enumerate.cpp:120-943).
 *   vx, vy : nel*4 CCW vertex coordinates of straight-sided quads
 *   nbr    : nel*4, for side s of element e the index e2*4+s2 of the side
 *            it is joined to (interior neighbour, or periodic partner: the
 *            role of semtex's shared NODES and <P> surfaces), -1 on a
 *            boundary.  Joined sides are traversed in opposite directions.
 *   skind  : nel*4 side kinds: 0 joined, 1 essential (<D>), 2 natural
 *            (<N>).  The boundary list is every side with skind != 0 in
 *            (element, side) order.
 *   svv_mn, svv_eps, kinvis : SVV tokens SVV_MN, SVV_EPSN, KINVIS
 *            (src/svv.cpp:41-79); svv_mn < 0 switches SVV off.
 * Host-only; device copies are made on first use.                         */
int semx_mesh_create (int np, int nel, const double* vx, const double* vy,
                      const int* nbr, const int* skind, int svv_mn,
                      double svv_eps, double kinvis, semx_mesh* mesh);
/* The same mesh numbered as semtex numbers it: btog and bmask (nel*next
 * each, element by element, sides 0..3 CCW, each side from its first corner
 * -- the bmap / mask columns of session.num), NGLOBAL and NSOLVE, as
 * BCmgr::buildnum reads them from the file that enumerate writes
 * (src/bcmgr.cpp:586-824, utility/enumerate.cpp:426-524).  With the same
 * numbering the band layout and the pinned node of a singular mode-0
 * system (src/matrix.cpp:167-169) are semtex's.  SEMX_E_MESH when the
 * numbering does not join exactly the element-boundary nodes the
 * connectivity joins, mask exactly the essential ones, or number the
 * unknowns below nsolve and the essential nodes from nsolve up.          */
int semx_mesh_create_numbered (int np, int nel, const double* vx,
                               const double* vy, const int* nbr,
                               const int* skind, int svv_mn, double svv_eps,
                               double kinvis, const int* btog,
                               const int* bmask, int nglobal, int nsolve,
                               semx_mesh* mesh);
/* The same mesh numbered as semtex's `enumerate -O<optlevel>` numbers it
 * when it writes session.num (utility/enumerate.cpp:540-943; BCmgr runs
 * `enumerate -O1` itself when the file is missing, src/bcmgr.cpp:620-632):
 * the traverse numbering of Mesh::buildMap (src/mesh.cpp:889-1015), the
 * unknowns before the essential nodes (Nsys::sortGid), and for optlevel
 * 1-3 a reverse Cuthill-McKee order of the unknowns (SPARSPAK FNROOT/RCM,
 * femlib/netlib.f) from the root that level picks -- 3, enumerate's
 * default and the regression scripts', tries every unknown as the root
 * (one RCM per unknown).  SEMX_E_ARG for optlevel outside 0-3,
 * SEMX_E_UNSUPPORTED if the unknowns' graph is disconnected.            */
int semx_mesh_create_enumerated (int np, int nel, const double* vx,
                                 const double* vy, const int* nbr,
                                 const int* skind, int svv_mn, double svv_eps,
                                 double kinvis, int optlevel, semx_mesh* mesh);
/* Cylindrical coordinates for the mesh (the CYLINDRICAL token; x axial,
 * y radial, every node at y >= 0): the geometric factors of the stiffness
 * terms and the side weights of natural BCs are multiplied by y, and every
 * Helmholtz system built on the mesh afterwards is the premultiplied one of
 * src/element.cpp:1316-1426 and 1464-1650 -- mass term Q4 y (lambda2 +
 * betak^2 / y^2), zero on the axis -- so the forcing handed to a solve is
 * y f (AuxField::mulY, elliptic/helmholtz.cpp:45-63), weighted by the
 * unweighted mass (Element::weight, src/element.cpp:831-837).  betak is
 * beta k as for the scalar fields c, p and u (Field::modeConstant,
 * src/field.cpp:1121-1146).  Axis sides (semtex's <A> BCs,
 * src/bcmgr.cpp:133-184) are given by the caller per mode range, as
 * semtex's per-mode numberings 'c' / 'C' hold them: natural with zero
 * values for mode 0, essential zero for modes >= 1 (a second mesh of the
 * same elements, and a system with base_mode 1 on it).  The DNS step does
 * not take cylindrical meshes (SEMX_E_UNSUPPORTED).  SEMX_E_MESH if a node
 * has y < 0; SEMX_E_ARG once a system has been built on the mesh.         */
int semx_mesh_set_cylindrical (semx_mesh mesh);
int semx_mesh_destroy (semx_mesh mesh);
/* info[10] = np nel next nint nglobal nsolve nband nbound nPlane planeSize,
 * nband including the diagonal (utility/enumerate.cpp:903-943).           */
int semx_mesh_info (semx_mesh mesh, int* info);
/* Node coordinates, nel*np*np each (host arrays).                          */
int semx_mesh_nodes (semx_mesh mesh, double* x, double* y);
/* Boundary list: element and side of each boundary (host arrays, nbound). */
int semx_mesh_bsides (semx_mesh mesh, int* belmt, int* bside);
/* NumberSys btog (nel*next), bmask (nel*next), emask (nel)
 * (src/numbersys.h:4-75).                                                  */
int semx_mesh_numbering (semx_mesh mesh, int* btog, int* bmask, int* emask);
/* Number of distinct element geometric-factor sets (semtex Family
 * semantics, femlib/family.c:171-192: equal within 6e-14).                */
int semx_mesh_nsets (semx_mesh mesh, int* nsets);

/* --- Helmholtz systems: ModalMatrixSys --------------------------------------
 * Replaces ModalMatrixSys::ModalMatrixSys (src/matrix.cpp:33-110) and
 * MatrixSys::MatrixSys (src/matrix.cpp:139-323): one system per Fourier
 * mode base_mode..base_mode+nmodes-1 with betak2 = (beta*mode)^2, for a
 * field of nz planes.  method: SEMX_DIRECT, SEMX_JACPCG or SEMX_MIXED.
 * Setup (elemental static condensation, band assembly and Cholesky
 * factorisation) runs on the host; factors are uploaded to the device.   */
int semx_helm_create (semx_mesh mesh, double lambda2, double beta,
                      int base_mode, int nmodes, int nz, int method,
                      semx_helm* helm);
/* The same with semtex's Fourier-direction spectral vanishing viscosity
 * (SVV_MZ, SVV_EPSZ, KINVIS): when lambda2 > EPSDP, betak2 of mode k is
 * weighted by 1 + SVV_EPSZ/KINVIS exp(-((N-k)/(SVV_MZ-k))^2) for k > SVV_MZ
 * (SVV::coeffs_z, src/svv.cpp:82-123, applied in ModalMatrixSys,
 * src/matrix.cpp:78-83), N = nmode_total - 1 with nmode_total the field's
 * number of Fourier modes (Geometry::nMode, src/geometry.h:42; it exceeds
 * nmodes when the system holds one process's share).  svv_mz < 0 or
 * SVV_EPSZ <= EPSDP leaves the weights at 1.                              */
/* The same for the field named `field` ('c', 'p', 'u', 'v' or 'w'), whose
 * name picks the Fourier constant as Field::modeConstant does
 * (src/field.cpp:1121-1146, used by ModalMatrixSys, src/matrix.cpp:60-63):
 * beta k, except on a cylindrical mesh for the coupled fields v~ ('v':
 * beta k + 1) and w~ ('w': beta k - 1), both 1 for mode 0.               */
int semx_helm_create_field (semx_mesh mesh, double lambda2, double beta,
                            int base_mode, int nmodes, int nz, int method,
                            int field, semx_helm* helm);
int semx_helm_create_svv (semx_mesh mesh, double lambda2, double beta,
                          int base_mode, int nmodes, int nz, int method,
                          int svv_mz, double svv_epsz, double kinvis,
                          int nmode_total, semx_helm* helm);
int semx_helm_destroy (semx_helm helm);
/* info[8] = nmodes nz method nsolve(mode0) nband nblk(block size)
 *           bytes_per_solve_factor(MB*1000, int) singular(mode0)           */
int semx_helm_info (semx_helm helm, int* info);
/* How the direct planes' band sweeps run; no counterpart in semtex (whose
 * dpbtrs is one call) -- and how the PCG planes run.  info[6] = partitions
 * per mode of the partitioned sweep (1: not available), the sweep variant
 * the first call picked for 1, 2 and 3 fields (0 before that call; 1-7 one
 * workgroup per mode, field or plane over the whole band, 8 the
 * partitioned sweep), workgroups per mode of the cooperative sweep (0: not
 * used), workgroups per PCG plane in the last solve.                      */
int semx_helm_sweep_info (semx_helm helm, int* info);
/* Host copies of one element's condensed matrices for mode index m
 * (HelmholtzSC, src/element.cpp:131-229): hbb next^2, hbi next*nint,
 * hii nint^2, all row-major in emap order.                                 */
int semx_helm_element (semx_helm helm, int m, int elmt, double* hbb,
                       double* hbi, double* hii);

/* --- the solve: Field::solve ------------------------------------------------
 * Replaces Field::solve (src/field.cpp:509-725) for a field of nz planes:
 * f (forcing, device, nz*planeSize) is destroyed, u (device, nz*planeSize)
 * holds the initial guess (used by JACPCG) and receives the solution;
 * bc (device, nz*nbound*np) holds the boundary values in Fourier space.
 * tol_rel, step_max: TOL_REL, STEP_MAX for JACPCG (field.cpp:649-706).
 * The call is asynchronous on stream unless iters (host) is non-NULL: then
 * it waits for the stream and stores the largest PCG iteration count
 * (summed over inner solves with SEMX_PCG_FP32_REFINED), and returns
 * SEMX_E_MAXITER when a PCG plane hit its limit (semtex prints a WARNING,
 * field.cpp:713, and keeps the last iterate; so does this call).          */
int semx_helm_solve (semx_helm helm, double* f, const double* bc, double* u,
                     double tol_rel, int step_max, int* iters, void* stream);

/* Waits for stream and reports whether any solve of this system since the
 * last check had a cooperative band sweep whose bounded hand-off wait
 * expired (SEMX_E_TIMEOUT; its solution is wrong), then clears the mark.
 * No counterpart in semtex, whose solves are synchronous.  A solve called
 * with iters != NULL and semx_dns_run make this check themselves, so a
 * timeout is reported by the first synchronising call after it, never
 * lost; callers that only issue asynchronous solves (iters = NULL, graph
 * replays) call this at their own synchronisation points.                */
int semx_helm_check (semx_helm helm, void* stream);

/* Per-plane JACPCG iteration counts of the last solve of the system (its
 * first field): iters[nz] (host), summed over inner solves with refinement,
 * negative where the step limit was hit, 0 for planes solved directly and
 * for the skipped Nyquist plane.  semtex reports the count per plane in its
 * verbose output (src/field.cpp:714-718).  Waits for stream.             */
int semx_helm_iters (semx_helm helm, int* iters, void* stream);

/* Field::solve for nfield (1..3) fields that share one ModalMatrixSys, as
 * the velocity components u, v, w do in semtex (the MatrixSys cache,
 * src/matrix.cpp:85-91, and Solve per component, dns/integrate.cpp:
 * 365-395).  f, bc, u: host arrays of nfield device pointers, each as in
 * semx_helm_solve (bc may be NULL on a mesh without boundaries).  The band
 * factor of each mode is streamed once for all fields' planes.            */
int semx_helm_solve_fields (semx_helm helm, int nfield, double* const* f,
                            const double* const* bc, double* const* u,
                            double tol_rel, int step_max, int* iters,
                            void* stream);

/* Precision of the JACPCG planes of a system.  SEMX_PCG_FP64 (the default)
 * is the FP64 loop of Field::solve (src/field.cpp:627-720).
 * SEMX_PCG_FP32_REFINED has no counterpart in semtex: the residual
 * r = b - A x is formed in FP64, A d = r is solved by FP32 Jacobi-PCG to
 * relative residual inner_tol, x += d, repeated (at most max_outer times)
 * until |r| <= tol_rel |b| in FP64 -- the same stopping test as FP64.
 * SEMX_PCG_BF16_REFINED is the same refinement with the inner iteration's
 * element operator (Element::HelmholtzOp, src/element.cpp:1621-1650) on the
 * BF16 matrix cores: operands rounded to BF16, FP32 accumulation; the FP64
 * residuals use the FP64 operator.  Its inner operator is accurate to about
 * 1e-2 relative, so inner_tol of order 1e-2 and a larger max_outer suit it.
 * SEMX_E_UNSUPPORTED for np > 16.                                         */
int semx_helm_set_pcg (semx_helm helm, int precision, double inner_tol,
                       int max_outer);

/* --- many independent 2D problems in one launch ----------------------------
 * No counterpart in semtex, which solves one problem per process: count
 * systems created with nz = 1, nmodes = 1 and method SEMX_JACPCG (each on
 * its own mesh, np, lambda2), all with the same semx_helm_set_pcg setting,
 * are solved together in one launch; the workgroups of a problem do all
 * of Field::solve's JACPCG branch (src/field.cpp:564-725) for its plane:
 * -(M f + H g) + <h,w>, x0 from u, the iteration, essential re-imposition.
 * A problem whose expected cost exceeds the batch's per-CU share gets
 * several workgroups (up to 8; SEMX_BATCH_WG=1: one each), launched first;
 * the rest one workgroup each, largest first.
 * f, bc, u: host arrays of count device pointers (bc[i] may be NULL when
 * mesh i has no boundary); iters as in semx_helm_solve.  The systems must
 * outlive the batch.                                                      */
int semx_batch_create (int count, const semx_helm* helms, semx_batch* batch);
int semx_batch_destroy (semx_batch batch);
/* The element operator each problem's inner iteration uses (opk[count]:
 * 0 block-wide scalar, 1 wavefront-level scalar, 2 BF16 matrix cores; -1
 * before the first solve) and the host time the batch tuner took (ms).
 * The first semx_batch_solve tunes: per class of problems with the same np
 * and element count it times each operator the precision allows on the
 * class's first problem (40 iterations, into a scratch solution) and keeps
 * the fastest per iteration; it synchronises the stream, so make it before
 * capturing a graph.  SEMX_BATCH_OP=block|wave|mfma forces one operator
 * (the BF16 batch offers wave and mfma).                                 */
int semx_batch_info (semx_batch batch, int* opk, double* tune_ms);
int semx_batch_solve (semx_batch batch, const double* const* f,
                      const double* const* bc, double* const* u, double tol_rel,
                      int step_max, int* iters, void* stream);

/* --- field operators (device data, caller's stream) ------------------------
 * dDFTr (femlib/fourier.c:47-171): ntrn real transforms of length tlen,
 * successive points ntrn apart (a field of tlen planes of ntrn values),
 * in place.  sign SEMX_FORWARD: X_j = (1/tlen) sum x_k exp(-2 pi i jk/tlen)
 * packed per column as [Re0, Re_tlen/2, Re1, Im1, Re2, Im2, ...];
 * SEMX_INVERSE: the unscaled inverse of that packing.  tlen must be
 * 2^p 3^q 5^r and ntrn even (SEMX_E_ARG otherwise, as dDFTr's error);
 * lengths above 256 return SEMX_E_UNSUPPORTED.
 * Uses a library scratch buffer: concurrent calls from several host
 * threads are not supported.                                              */
int semx_dft_r (double* data, int tlen, int ntrn, int sign, void* stream);
/* The same transform, allocating nothing: for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in semtex.  It works in a static
 * device array of the library (the transform matrix and a staging copy of
 * the field, about 48.5 MiB, held on every device the library runs on until
 * the process ends) and otherwise behaves as semx_dft_r.  Limits: tlen <= 256
 * and tlen*ntrn <= SEMX_DFT_NOALLOC_MAX; larger requests return
 * SEMX_E_SCRATCH.  Every call on one device shares that device's array:
 * calls that overlap on one device (several streams or host threads) race.
 * The array's device address is looked up by semx_device_init (call it once
 * per device before capturing), else by the first call on a device.
 * Results agree with semx_dft_r to rounding (its transform matrix is formed
 * on the device in FP64 rather than on the host in long double).           */
#define SEMX_DFT_NOALLOC_MAX ((size_t)6 << 20)   /* doubles (48 MiB)        */
int semx_dft_r_noalloc (double* data, int tlen, int ntrn, int sign, void* stream);
/* AuxField::gradient (src/auxfield.cpp:514-665) on nz planes of data (in
 * place): dir 0 = d/dx, 1 = d/dy (Element::grad, src/element.cpp:552-597,
 * per plane), 2 = d/dz of Fourier planes (mode k: Re <- -beta k Im,
 * Im <- beta k Re; planes 0 and 1 zeroed).                                 */
int semx_gradient (semx_mesh mesh, int nz, double beta, double* data, int dir,
                   void* stream);
/* Field::smooth (src/field.cpp:190-236): mass-weighted average of the
 * element-boundary values of every plane.                                  */
int semx_smooth (semx_mesh mesh, int nz, double* data, void* stream);
/* AuxField::couple and Field::coupleBCs (src/auxfield.cpp:1352-1420,
 * src/field.cpp:1043-1118): the radial and azimuthal fields of a
 * cylindrical 3D run coupled for their viscous solves, per Fourier mode k
 * (planes 2k, 2k+1 of n values each: planeSize for fields, nbound*np for
 * BC lines) -- dir = 1: v~ = v + i w, w~ = v - i w; dir = -1: v = (v~ +
 * w~)/2, w = i (w~ - v~)/2.  Mode 0 is left alone when base_mode is 0 (it
 * has no coupling, semtex's kLo), as are 2D fields (nz = 1).  Device
 * pointers, in place.                                                     */
int semx_couple (double* v, double* w, int nz, int n, int base_mode, int dir,
                 void* stream);

/* --- DNS time stepping: integrate -------------------------------------------
 * Replaces the time loop of integrate (dns/integrate.cpp:79-221) with the
 * skew-symmetric (advection 0, nonlinear.cpp:36-211), alternating
 * (advection 1, :214-411; dns's default), convective (advection 2,
 * :414-525) or Stokes (advection 3, :528-570: none) nonlinear terms -- dns
 * -S, default, -C, -N (dns/drive.cpp:114-158) -- waveProp, setPForce, project and
 * the pressure / viscous Field::solve calls (DIRECT; JACPCG start-up
 * systems while step < norder = N_TIME), Cartesian, on a mesh without
 * boundaries (periodic in x and y; SEMX_E_UNSUPPORTED otherwise: a mesh
 * with boundaries is set up by semx_dns_create_bc).  nz = 1
 * is 2D (u, v); nz > 1 is 3D3C with Fourier z (nz = 2 or 2^p 3^q 5^r).
 * create builds the velocity (lambda2 = alpha0/(dt kinvis)) and pressure
 * (lambda2 = 0) systems and the device history (Us, Uf of integrate).    */
int semx_dns_create (semx_mesh mesh, int nz, double beta, double dt,
                     double kinvis, int norder, int advection, double tol_rel,
                     int step_max, semx_dns* dns);
/* The same on a mesh with boundaries (dns/dns.cpp:55-133 with BCmgr,
 * src/bcmgr.cpp): vmesh numbers the velocity, pmesh the pressure -- two
 * numberings of one element mesh with one boundary list (the same
 * vertices, neighbours and sides; SEMX_E_MESH otherwise), as semtex's
 * session.num holds one numbering per field mask.  bcu, bcv, bcw (device,
 * nz*nbound*np each in Fourier space, as semx_helm_solve's bc; bcw NULL
 * when nz = 1) are the velocity BC lines: values on essential sides,
 * fluxes on natural ones, fixed in time; they are copied.  The pressure
 * takes the computed high-order flux on its natural sides (KIO91:
 * BCmgr::maintainFourier and evaluateCNBCp, src/bcmgr.cpp:1039-1337, with
 * u.n and flux levels kept across steps) and p = 0 on its essential ones. */
int semx_dns_create_bc (semx_mesh vmesh, semx_mesh pmesh, int nz, double beta,
                        double dt, double kinvis, int norder, int advection,
                        double tol_rel, int step_max, const double* bcu,
                        const double* bcv, const double* bcw, semx_dns* dns);
/* The ITERATIVE token (dns -i / -ii, dns/integrate.cpp:317-362 preSolve):
 * level 1 solves the velocity of the steady steps with Jacobi-PCG
 * (JACPCG) instead of DIRECT, level 2 also the pressure with MIXED (mode 0
 * DIRECT, the other modes JACPCG); 0 restores DIRECT.  The PCG solves stop
 * at the DNS's tol_rel and step_max.  Only before the first step
 * (SEMX_E_ARG otherwise, or for a level outside 0..2).                    */
int semx_dns_set_iterative (semx_dns dns, int level);
/* New velocity BC lines (device, nz*nline each, as semx_dns_create_bc takes
 * them) for the steps that follow: the values of time-dependent BCs, which
 * semtex re-evaluates after advancing the time of each step and before its
 * viscous solve (Field::evaluateBoundaries, dns/integrate.cpp:196-201), so
 * a caller stepping one step at a time hands over the lines of that step's
 * end time before running it.  Copied on the stream; graphs captured for
 * the DNS keep reading the same arrays.  SEMX_E_ARG for a DNS created
 * without boundaries.                                                      */
int semx_dns_set_bc (semx_dns dns, const double* bcu, const double* bcv,
                     const double* bcw, void* stream);
int semx_dns_destroy (semx_dns dns);
/* nstep steps; u, v, w (NULL in 2D), p: device fields of nz*planeSize in
 * Fourier space, read at entry and written at exit.  The history carries
 * over between calls, as integrate's static storage does.  Returns after
 * the stream has finished the steps, with SEMX_E_TIMEOUT if a cooperative
 * band sweep among them timed out (semx_helm_check).                     */
int semx_dns_run (semx_dns dns, int nstep, double* u, double* v, double* w,
                  double* p, void* stream);
/* info[4] = steps taken, nz, velocity components, norder                  */
int semx_dns_info (semx_dns dns, int* info);
/* Replay steady time steps of semx_dns_run from HIP graphs (on != 0) or
 * launch every kernel (0, the default).  No counterpart in semtex, whose
 * time loop (dns/integrate.cpp:148-219) calls its operators one by one.
 * A step permutes the roles of its buffers (transform output and spare,
 * swapData, the rolled time levels), so a graph is captured, on first use,
 * for each role assignment the steps meet (a cycle of 88 steps for N_TIME =
 * 3 with the skew-symmetric terms, 182 with the alternating form) and
 * replayed whenever that assignment recurs.  Start-up steps, the first
 * steady step and steps taken while kernel timing is on run eagerly.
 * Replayed steps equal eager ones bit for bit.                           */
int semx_dns_set_graph (semx_dns dns, int on);
/* Graphs captured so far and steps replayed from them.                     */
int semx_dns_graph_info (semx_dns dns, int* ngraphs, long long* replays);

/* --- field files ------------------------------------------------------------
 * semtex's field-file format (.fld, .rst, .chk): writeField
 * (src/misc.cpp:106-180) and Domain's stream input (src/domain.cpp:
 * 329-398).  A dump is a 10-line header (Session, Created, "Nr, Ns, Nz,
 * Elements", Step, Time, Time step, Kinvis, Beta, Fields written, Format;
 * each value left-justified in 25 columns) followed, for each field named,
 * by its nz physical-space planes of nPlane doubles in the writer's byte
 * order.
 * semx_field_write: nfield device fields in Fourier space (nz*planeSize
 * each, named by the letters of names) written as one dump, each
 * transformed to physical space on a scratch copy with the Nyquist plane
 * zeroed first (Domain::dump, src/domain.cpp:188-248, and
 * Domain::transform, :250-264); append != 0 adds the dump to the end of
 * the file (a .fld written every IO_FLD steps).
 * semx_field_read: dump number `dump` (0 = first) of a file read into the
 * device fields whose names the file holds (others are left as they are;
 * fields of the file not asked for are skipped), transformed forward with
 * the Nyquist plane zeroed (Domain::restart, src/domain.cpp:152-180);
 * byte-swapped when the file's byte order differs.  step, time (may be
 * NULL) receive the header's values, fields_in_file (may be NULL, 64
 * chars) its field names.  SEMX_E_MESH when np, nz or the element count
 * differ from the mesh's (readField, src/misc.cpp:196-205).               */
int semx_field_write (const char* path, int append, semx_mesh mesh, int nz,
                      int nfield, const char* names, double* const* fields,
                      const char* session, int step, double time, double dt,
                      double kinvis, double beta, void* stream);
int semx_field_read (const char* path, int dump, semx_mesh mesh, int nz,
                     int nfield, const char* names, double* const* fields,
                     int* step, double* time, char* fields_in_file,
                     void* stream);

/* --- kernel timing -------------------------------------------------------
 * No counterpart in semtex (dclock(), veclib/util.c:65-85, is never called
 * on the solve path).  When enabled, each kernel launch of the solve path is
 * bracketed by HIP events on its own stream; query returns the summed
 * device time (ms) and launch count per kernel name (e.g. "k_band_solve").*/
int semx_profile_enable (int on);
int semx_profile_reset (void);
int semx_profile_query (const char* kernel, double* total_ms, long long* count);

/* --- device memory helpers (plumbing for callers without their own) ---- */
int semx_malloc (void** ptr, size_t bytes);
int semx_free (void* ptr);
int semx_memcpy_htod (void* dst, const void* src, size_t bytes);
int semx_memcpy_dtoh (void* dst, const void* src, size_t bytes);
int semx_synchronize (void);

#ifdef __cplusplus
}
#endif
#endif
