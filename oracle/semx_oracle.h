/*
 * semx_oracle.h -- CPU FP64 restatement of semtex's spectral-element /
 * Fourier Helmholtz and DNS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (semtex_amd/libsemx.so) never links or calls it.
 *
 * Every function names the reference file:line it restates.  Storage
 * conventions follow semtex:
 *   - element data row-major, np x np, data[e*np*np + i*np + j], i = s
 *     (row), j = r (column)                         (src/element.h:23-24)
 *   - planes of nel*np*np doubles, plane stride planeSize = nPlane
 *     rounded up to even                       (src/geometry.cpp:105-108)
 *   - Fourier planes: 2m = Re(mode m), 2m+1 = Im(mode m), plane 1 holds
 *     the Nyquist datum                            (src/auxfield.h:18-29)
 *   - global numbering: element-boundary nodes first, essential-BC nodes
 *     last, then element interiors in emap order  (src/field.cpp:541-545)
 */
#ifndef SEMX_ORACLE_H
#define SEMX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* --- quadrature & 1D operators ------------------------------------------ */
void orc_zwglj (double* z, double* w, int np);                 /* polylib.c:312-340 */
void orc_dglj  (double* D, double* DT, const double* z, int np); /* polylib.c:513-556 */
void orc_legtran (int np, double* FW, double* BW);              /* operators.c:391-480 */
void orc_svv_operators (int np, int svv_mn, double svv_eps, double kinvis,
                        double* SDV, double* SDT);              /* svv.cpp:41-79,126-185 */
void orc_svv_coeffs_z (int nummodes, int nmode, int base, int mz, double eps, double nu,
                       double* S);                            /* svv.cpp:82-123 */
void orc_set_svv_z (int mz, double eps, double nu, int nmode); /* SVV_MZ, SVV_EPSZ */
void orc_edgemaps (int np, int* emap, int* pmap);               /* mapping.c:46-132 */

/* --- mesh, geometry, numbering ------------------------------------------ */
typedef struct orc_mesh orc_mesh;

/* Quad mesh with straight sides.  vx,vy: nel*4 CCW vertex coordinates;
 * nbr: nel*4, the side (e2*4+s2) each side is joined to (neighbour or
 * periodic partner), -1 on a boundary;
 * skind: nel*4 side kinds, 0 = joined, 1 = essential (<D>),
 * 2 = natural (<N>).  The boundary list is every side with skind != 0 in
 * (element, side) order; BC lines are nbound*np values per plane in that
 * order, each side traversed CCW.                                        */
orc_mesh* orc_mesh_create (int np, int nel, const double* vx, const double* vy,
                           const int* nbr, const int* skind,
                           int svv_mn, double svv_eps, double kinvis);
void orc_mesh_destroy (orc_mesh* M);
/* info: np nel next nint nglobal nsolve nband nbound nPlane planeSize    */
void orc_mesh_info (const orc_mesh* M, int* info);
void orc_mesh_nodes (const orc_mesh* M, double* x, double* y);
void orc_mesh_bsides (const orc_mesh* M, int* belmt, int* bside);
void orc_mesh_numbering (const orc_mesh* M, int* btog, int* bmask, int* emask);
int  orc_mesh_renumber (orc_mesh* M, const int* btog, const int* bmask, int nglobal,
                        int nsolve);                /* src/bcmgr.cpp:586-824 */
int  orc_mesh_enumerate (orc_mesh* M, int optlevel);   /* utility/enumerate.cpp */
/* CYLINDRICAL = 1: x axial, y radial; 1 if a node has y < 0 */
int  orc_mesh_set_cylindrical (orc_mesh* M);
/* Field name for Field::modeConstant in the field solves ('c' default) */
void orc_set_field_name (int name);
/* AuxField::couple / Field::coupleBCs: dir 1 forward, -1 inverse */
void orc_couple (double* v, double* w, int nz, int n, int base, int dir);
void orc_mesh_factors (const orc_mesh* M, double* Q1, double* Q2, double* Q3,
                       double* Q4, double* drdx, double* dsdx, double* drdy,
                       double* dsdy, int* null_flags /* nel*5 */);
void orc_mesh_operators (const orc_mesh* M, double* z, double* w, double* DV,
                         double* SDV, int* emap);
void orc_mesh_imass (const orc_mesh* M, double* imass);
void orc_mesh_barea (const orc_mesh* M, double* barea /* nbound*np */);

/* --- element operators --------------------------------------------------- */
/* Element::HelmholtzOp  (element.cpp:1621-1650)                           */
void orc_helm_op (const orc_mesh* M, int e, double lambda2, double betak2,
                  const double* src, double* tgt);
/* Full np^2 x np^2 elemental matrix from Element::HelmholtzRow
 * (element.cpp:1439-1509), row-major                                      */
void orc_helm_matrix (const orc_mesh* M, int e, double lambda2, double betak2,
                      double* H);
/* Element::HelmholtzSC (element.cpp:131-229)                              */
void orc_helm_sc (const orc_mesh* M, int e, double lambda2, double betak2,
                  double* hbb, double* hbi, double* hii);
/* Element::HelmholtzDiag (element.cpp:1512-1561), emap order              */
void orc_helm_diag (const orc_mesh* M, int e, double lambda2, double betak2,
                    double* diag);

/* --- matrix systems & solves --------------------------------------------- */
enum { ORC_DIRECT = 0, ORC_JACPCG = 1 };
typedef struct orc_sys orc_sys;
/* MatrixSys (matrix.cpp:139-323)                                          */
orc_sys* orc_sys_create (const orc_mesh* M, double lambda2, double betak2,
                         int method);
void orc_sys_destroy (orc_sys* S);
/* info: nsolve (after singular pin), nband, singular                      */
void orc_sys_info (const orc_sys* S, int* info);
/* Band factor in LAPACK 'U' storage, nband*nsolve                         */
void orc_sys_band (const orc_sys* S, double* H);
/* Field::solve for one plane (field.cpp:509-725).  forcing is destroyed;
 * unknown holds the initial guess (JACPCG) and receives the solution.
 * Returns PCG iteration count (0 for DIRECT), -1 on step-limit warning. */
int orc_solve_plane (const orc_sys* S, double* forcing, const double* bc,
                     double* unknown, double tol_rel, int step_max);
/* Global Helmholtz operator (field.cpp:771-856) on globally numbered x.  */
void orc_helm_global (const orc_sys* S, const double* x, double* y);

/* Field::solve over all planes of a 2D x Fourier field with ModalMatrixSys
 * semantics (matrix.cpp:33-110): plane k uses mode k>>1, betak2 =
 * (beta*mode)^2, plane 1 (Nyquist) skipped; identical systems shared.
 * f, u: nz*planeSize; bc: nz*nbound*np.  nthreads: OpenMP threads.
 * The solve is repeated nrep times (forcing restored between repeats);
 * t_solve receives the mean seconds per all-plane solve, excluding setup.*/
int orc_field_solve (const orc_mesh* M, int nz, double lambda2, double beta,
                     int method, double* f, const double* bc, double* u,
                     double tol_rel, int step_max, int nthreads, int nrep,
                     double* t_solve);
/* The same, one repetition, with each plane's PCG iteration count in
 * iters[nz] (-1: step limit; plane 1 untouched when nz > 1).              */
int orc_field_solve_iters (const orc_mesh* M, int nz, double lambda2, double beta,
                           int method, double* f, const double* bc, double* u,
                           double tol_rel, int step_max, int* iters);
/* The same for the planes of modes base..base+nz/2-1 (one MPI process's
 * share in semtex, Geometry::baseMode): plane k uses mode base + k>>1 and
 * plane 1 is solved unless base == 0 (field.cpp:566 skips it on the root
 * process only).                                                          */
int orc_field_solve_base (const orc_mesh* M, int nz, double lambda2, double beta,
                          int base, int method, double* f, const double* bc, double* u,
                          double tol_rel, int step_max, int* iters);

/* --- Fourier transform & field operators -------------------------------- */
/* dDFTr (fourier.c:47-171) semantics: ntrn lines of tlen reals strided by
 * ntrn; sign +1 forward (scaled 1/tlen), -1 inverse.  Returns 0 on
 * success, -1 if tlen has prime factors other than 2,3,5 or ntrn is odd. */
int orc_dftr (double* data, int tlen, int ntrn, int sign);
/* AuxField::gradient (auxfield.cpp:514-665) on nz planes                  */
void orc_gradient (const orc_mesh* M, int nz, double beta, double* data,
                   int dir);
/* Field::smooth (field.cpp:190-236) on nz planes                          */
void orc_smooth (const orc_mesh* M, int nz, double* data);

/* --- DNS time step (dns/integrate.cpp:79-221), Cartesian, no BCs -------- */
typedef struct orc_dns orc_dns;
/* u,v,w,p in Fourier space, each nz*planeSize.  advection: 0 skew-symmetric,
 * 1 alternating skew-symmetric (nonlinear.cpp:36-211, 214-411).           */
orc_dns* orc_dns_create (const orc_mesh* M, int nz, double beta, double dt,
                         double kinvis, int norder, int advection,
                         double tol_rel, int step_max, int nthreads);
/* the same on a mesh with boundaries: M numbers the velocity (given
 * Fourier-space BC lines on its essential sides), Mp the pressure
 * (computed high-order flux on its natural sides, BCmgr::maintainFourier
 * and evaluateCNBCp, src/bcmgr.cpp:1039-1337)                          */
orc_dns* orc_dns_create_bc (const orc_mesh* M, const orc_mesh* Mp, int nz, double beta,
                            double dt, double kinvis, int norder, int advection,
                            double tol_rel, int step_max, int nthreads, const double* bcu,
                            const double* bcv, const double* bcw);
void orc_dns_set_bc (orc_dns* D, const double* bcu, const double* bcv, const double* bcw);
int  orc_dns_set_iterative (orc_dns* D, int level);   /* ITERATIVE 0-2 */
void orc_dns_destroy (orc_dns* D);
/* Advance nstep steps; fields are read from/written to u,v,w,p (ncom=3)
 * or u,v,p (ncom=2, pass w = NULL).                                       */
int orc_dns_run (orc_dns* D, int nstep, double* u, double* v, double* w,
                 double* p, double* t_step);

#ifdef __cplusplus
}
#endif
#endif
