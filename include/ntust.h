/*
 * ntust.h -- C ABI of libntust_hip.so, the gfx950 implementation of the
 * per-time-step hot path of the HSUAN221/NTUSTCFDLAB projection solver.
 *
 * Conventions
 *  - Every array argument is the host base address of a Fortran module array
 *    of the solver, i.e. its first element: (-1,-1,-1) of a halo array
 *    dimensioned (-1:nx+2,-1:ny+2,-1:nz+2), (1,1,1) of an interior array
 *    (1:nx,1:ny,1:nz), element -1 of a metric array (-1:n+2).  Column-major,
 *    x fastest, real*8.  From Fortran these bind as `real(c_double) :: a(*)`
 *    and are passed without a copy (variables_module.f90:159,215-225,281-316).
 *  - Scalars are passed by value.  istart/iend are the z-slab bounds of
 *    Mpi_division.f90:11-24; on one GPU they are 1 and nz.
 *  - Every call returns an int status: 0 ok, > 0 a warning (NTUST_WARN_*),
 *    < 0 an error (NTUST_E_*; ntust_last_error() has the message).  The
 *    reference has no error convention: its callers ignore ierr and a
 *    non-converged pressure solve continues silently (GS.f90:11), which is
 *    what ignoring a warning does.
 *  - Not re-entrant: one host thread, outside any OpenMP parallel region,
 *    as main.f90:139-355 calls the hot path.  Calls are stream ordered and
 *    synchronous at return.
 *  - Drop-in calls (ntust_<subroutine>) copy the subroutine's read set to
 *    the device, run it and copy its write set back, so each replaces one
 *    reference call without changing the caller.  The device-resident calls
 *    (ntust_state_*, ntust_step) keep the state in HBM across steps.
 */
#ifndef NTUST_H
#define NTUST_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NTUST_OK = 0,
    NTUST_WARN_ITMAX = 1,        /* SOR stopped at itmax before max|dp| <= zeta */
    NTUST_E_NOT_INIT = -1,
    NTUST_E_ARG = -2,
    NTUST_E_HIP = -3,
    NTUST_E_ALLOC = -4,
    NTUST_E_STATE = -5,          /* a recomputation did not reproduce the forward pass */
    NTUST_E_IO = -6              /* a file could not be written or read, or is malformed */
};

/* pressure-solver orderings (ntust_configure sor_mode) */
enum {
    NTUST_SOR_LEXICOGRAPHIC = 0,  /* GS.f90 order, bit-identical, pipelined over sweeps */
    NTUST_SOR_REDBLACK = 1,       /* red-black, FP64 */
    NTUST_SOR_MIXED = 2           /* red-black, FP32 sweeps inside FP64 defect correction */
};

/* arithmetic of the resident step's QUICK predictor and eddy viscosity
 * (ntust_set_arithmetic); the pressure solve, projection, forcing and force
 * integral are unaffected */
enum {
    NTUST_ARITH_EXACT = 0,        /* the reference's operation order, bit-identical (default) */
    NTUST_ARITH_FAST = 1,         /* FP64, reciprocal metric tables, fused multiply-adds */
    NTUST_ARITH_FP32 = 2          /* the increment u* - u and nu_t in FP32 from FP64
                                     differences; u* = u + increment in FP64 */
};

/* advection schemes (ntust_configure scheme) */
enum {
    NTUST_QUICK_CENTRE = 0,       /* discretisation_QUICK_centre.f90 */
    NTUST_UPWIND = 1              /* discretisation_upwind.f90 */
};

/* ---- lifecycle ------------------------------------------------------ */

/* Sets up the device for an nx*ny*nz grid and copies the six metric
 * arrays the gridder built (gridder.f90:267-327; variables_module.f90:306-316).
 * Replaces the device-side part of the initialisation main.f90:77-92.
 * Calling it again re-initialises (finalize first). */
int ntust_init(int nx, int ny, int nz, const double *iDx, const double *Dxs,
               const double *iDy, const double *Dys, const double *iDz,
               const double *Dzs);

/* Frees all device memory.  Called before MPI_FINALIZE (main.f90:362). */
int ntust_finalize(void);

/* Text of the last error. */
const char *ntust_last_error(void);

/* The 18 boundary-condition strings of variables_module.f90:110-132 as
 * integer codes (0 'no-slip', 1 'Neumann', 2 'Dirichlet', else no rule),
 * order W_u W_v W_w E_u E_v E_w S_u S_v S_w N_u N_v N_w B_u B_v B_w F_u F_v F_w. */
int ntust_set_bc(const int *codes);

/* ---- drop-in replacements, one per reference subroutine -------------- */

/* smagorinsky.f90:42-87 GradientPhiGauss(Phi,dPhidX,Dxm,Dym,Dzm).  Phi is
 * read through the (nx,ny,nz) shape the reference declares; prebase (nx*ny
 * words, may be NULL = zeros) stands for the memory before Phi that the
 * k=1 plane reads.  dPhidX is (nx,ny,nz,3). */
int ntust_gradient_phi_gauss(const double *Phi, const double *prebase, double *dPhidX,
                             const double *Dxm, const double *Dym, const double *Dzm,
                             int istart, int iend);

/* smagorinsky.f90:1-37 CalculateSmagorinskyViscosity.  Cs is the model
 * constant (variables_module.f90:277).  dudx/dvdx/dwdx (nx,ny,nz,3) may be
 * NULL when the caller does not keep them.  pre_u/pre_v/pre_w: as prebase. */
int ntust_smagorinsky(const double *u, const double *v, const double *w,
                      const double *pre_u, const double *pre_v, const double *pre_w,
                      double Cs, int istart, int iend, double *dudx, double *dvdx,
                      double *dwdx, double *Viseff);

/* discretisation_QUICK_centre.f90:1-345.  LES is the integer switch of
 * main.f90:60.  Writes u_star, v_star, w_star over 1..nx,1..ny,istart..iend. */
int ntust_discretisation_quick_centre(const double *u, const double *v, const double *w,
                                      const double *Viseff, double dt, double nu, int LES,
                                      int istart, int iend, double *u_star, double *v_star,
                                      double *w_star);

/* discretisation_upwind.f90:1-275 with its undeclared switch Q1 = 0. */
int ntust_discretisation_upwind(const double *u, const double *v, const double *w,
                                double dt, double nu, double *u_star, double *v_star,
                                double *w_star);

/* GS.f90:1-106 gauss_seidel.  itmax is real*8 as in variables_module.f90:233.
 * Uses the ordering set by ntust_configure (default lexicographic).  Writes
 * p (interior), pre, *ik and *pChangeMax.  Returns NTUST_WARN_ITMAX when the
 * sweep cap ended the solve. */
int ntust_gauss_seidel(const double *u_star, const double *v_star, const double *w_star,
                       double *p, double *pre, double dt, double omega, double zeta,
                       double itmax, int *ik, double *pChangeMax);

/* calcul_new_velocity.f90:1-184.  u2, v2, w2 are read on their lowest ghost
 * layer (the centring of :171-181) and written on the interior. */
int ntust_calcul_new_velocity(const double *u_star, const double *v_star,
                              const double *w_star, const double *p, const double *ETA,
                              double u_solid, double v_solid, double w_solid, double dt,
                              int istart, int iend, double *u1, double *v1, double *w1,
                              double *u2, double *v2, double *w2, double *FX, double *FY,
                              double *FZ, double *uc, double *vc, double *wc);

/* calcul_new_velocity.f90:195-243 Updating_velocity: u = u2 on the interior. */
int ntust_updating_velocity(const double *u2, const double *v2, const double *w2,
                            int istart, int iend, double *u, double *v, double *w);

/* virtualForceIntegrator.f90:3-85.  lx is the span (variables_module.f90:19). */
int ntust_virtual_force_integrator(const double *FY, const double *FZ, double lx,
                                   double *cDrag, double *cLift);

/* No counterpart in the reference.  ntust_virtual_force_integrator for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): it calls no
 * hipMalloc, hipFree or allocating library routine.  FY and FZ are DEVICE
 * halo arrays of the grid of ntust_init (unlike every other entry point),
 * coef a DEVICE array of 2 doubles that gets cDrag, cLift.  The line sums go
 * to a static device scratch array of 2^20 doubles (8 MiB, held by each GPU
 * the library runs on until the process ends): grids with
 * 2(nx+2)(ny+3) > 2^20 are refused (NTUST_E_ARG).  Calls on one GPU share
 * that scratch, so calls that overlap on one GPU race; each call clears the
 * entries it needs zero.  Stream ordered on the library stream and
 * synchronous at return, like the other entry points. */
int ntust_virtual_force_integrator_noalloc(const double *FY, const double *FZ, double lx,
                                           double *coef);

/* boundary_conditions.f90:1-376 with the codes of ntust_set_bc: velocity
 * faces, whole-array copies into u1/u2/u_star (v, w alike; any target may
 * be NULL) and the Neumann pressure ghosts. */
int ntust_boundary_conditions(double *u, double *v, double *w, double *p, double *u1,
                              double *v1, double *w1, double *u2, double *v2, double *w2,
                              double *u_star, double *v_star, double *w_star);

/* ---- members of the Krylov pressure path ------------------------------
 * The reference's BiCGSTAB route (solver.f90, called from no time step of
 * main.f90) builds its system with these; they are provided for callers of
 * that route.  Interior arrays in the reference's x-fastest order. */

/* solver_1st_member.f90:1-79: the 7-diagonal matrix coef(ndim,7), ndim =
 * nx*ny*nz, column d holding diagonal d (1 centre, 2/3 +-x, 4/5 +-y, 6/7 +-z)
 * with the reference's boundary-row edits, and the offsets jcoef(4) =
 * (0, 1, nx, nx*ny). */
int ntust_poisson_coeffs(double *coef, int *jcoef);

/* The right-hand side b = mChange/dt of the pressure equation in the
 * face-area form of GS.f90:39-41,49 (interior array).  The reference's
 * solver_2nd_member.f90 (a volume-normalised variant with boundary pressure
 * terms) does not compile; this is the form the solve uses. */
int ntust_poisson_rhs(const double *u_star, const double *v_star, const double *w_star,
                      double dt, double *b);

/* solver_pressure_field.f90:1-21: pre = x and p = pre on the interior (x in
 * the order of the unknowns, one rank). */
int ntust_pressure_field(const double *x, double *p, double *pre);

/* matmul_ell (solver.f90:146-190): x = A y for the DIA matrix of
 * solver_1st_member, coef(n,7) (column-major) and jcoef(m), m <= 4
 * (jcoef(2..4) the offsets 1, nx, nx*ny).  Each row's terms are added in the
 * reference's order (diagonal, +1, -1, +nx, -nx, +nxny, -nxny): bit-identical
 * to a serial evaluation. */
int ntust_matmul_ell(double *x, const double *coef, const int *jcoef, const double *y, int n,
                     int m);

/* solver_BICG3 (solver.f90:1-144): unpreconditioned BiCGSTAB for A x = b
 * with the same matrix arguments, from the initial x, while
 * ||r|| / ||b|| > eps and fewer than itmax iterations.  x is overwritten by
 * the solution; *iterations and *norm get the final j and ||r|| / ||b||
 * (what the reference prints).  The work vectors p, r, r2, v, ss, t and
 * the MPI arguments of the reference are internal here.  Inner products
 * are deterministic tree sums, so iterates are reproducible run to run
 * but not bit-identical to the serial loop. */
int ntust_solver_bicg3(const double *coef, const int *jcoef, const double *b, double *x, int n,
                       int m, double eps, int itmax, int *iterations, double *norm);

/* ---- device-resident time stepping ------------------------------------ */

/* Step parameters (variables_module.f90:9-23,277; main.f90:44-60):
 * dt, nu = 1/Re, Cs, LES switch, advection scheme, SOR omega/zeta/itmax and
 * ordering, span lx for the force coefficients, force_every: evaluate the
 * force integral every force_every steps (0 = never; the reference: 1). */
int ntust_configure(double dt, double nu, double Cs, int LES, int scheme, double omega,
                    double zeta, double itmax, int sor_mode, double lx, int force_every);

/* Copies the state u, v, w, p (halo arrays, after boundary_conditions) and
 * ETA to the device. */
int ntust_state_upload(const double *u, const double *v, const double *w, const double *p,
                       const double *ETA);

/* Copies the state back; uc/vc/wc (cell-centred velocity of the last step)
 * and pre may be NULL. */
int ntust_state_download(double *u, double *v, double *w, double *p, double *pre,
                         double *uc, double *vc, double *wc);

/* Runs nsteps time steps in the order of main.f90:209-259 without leaving
 * the device.  Per-step outputs (host arrays of nsteps entries, each may be
 * NULL): SOR sweep count, last max|dp|, cDrag, cLift.  Returns
 * NTUST_WARN_ITMAX if any step hit the sweep cap. */
int ntust_step(int nsteps, int *ik, double *pchange, double *cdrag, double *clift);

/* ---- output, steady check and restart (resident state) ----------------- */

/* filereachtime (filer.f90:161-187): packs Qout(nx,ny,nz,5) = real*4 of
 * pre, uc, vc, wc, ETA on the device and copies it (20 B per cell) into
 * qout (the caller's real*4 array, or a library buffer when NULL).  With a
 * path, appends the four records nblocks; nx ny nz; temp x4 (1.0,
 * variables_module.f90:363); Qout as Fortran unformatted sequential records
 * (4-byte length before and after each), byte for byte what the
 * reference's write statements produce.  The caller names the file
 * (the reference: num in I4.4 plus '.q') and counts num. */
int ntust_output_q(const char *path, float *qout);

/* Reads a one-block solution file of this grid into qout (the reading
 * half of reading_variables, reading_data.f90:47-53).  Unlike the
 * reference, which reads inx, iny, inz and ignores them, a file of another
 * grid is refused (NTUST_E_ARG). */
int ntust_read_q(const char *path, float *qout);

/* reading_variables (reading_data.f90:56-83) on the resident state: u, v, w
 * = 0.5*(Q(i)+Q(i+1)) in single precision for i, j, k = 1..n-1, p and pre =
 * Q1, uc/vc/wc re-centred; then the boundary conditions, the call that
 * follows it (main.f90:96-100).  The state must have been uploaded
 * (ntust_state_upload): entries the reference leaves alone keep it. */
int ntust_reading_variables(const float *qout);

/* mesh.x of gridder.f90:355-378 from the cell centres Xs(1:nx), Ys(1:ny),
 * Zs(1:nz) (variables_module.f90:319-323); host only. */
int ntust_write_mesh_x(const char *path, const double *Xs, const double *Ys, const double *Zs);

/* check_steady (check_steady.f90:1-30) on the resident wc: when istep !=
 * isto, *velocity_difference = max |wc - last_velocity| over the interior
 * (otherwise it is left as it is); then last_velocity = wc.  last_velocity
 * is held on the device, zero until the first call, as the module array. */
int ntust_check_steady(int istep, int isto, double *velocity_difference);

/* Exact FP64 restart, replacing the single-precision .q restart of
 * reading_variables: writes the resident u, v, w, p, ETA (halo), pre, uc,
 * vc, wc and last_velocity (interior) with a header (grid, istep, time)
 * and an FNV-1a hash of the payload.  A run continued from the file is
 * bit-identical to the uninterrupted run.  Reading refuses another grid
 * (NTUST_E_ARG) and a damaged file (NTUST_E_IO), in which case the
 * resident state is unchanged. */
int ntust_restart_write(const char *path, int istep, double time);
int ntust_restart_read(const char *path, int *istep, double *time);

/* ---- airfoil geometry ---------------------------------------------------- */

/* RayCasting (airfoil.f90:580-892): the airfoil's volume fraction ETA.
 * az0, ay0: the polygon as the NACA table holds it (chord along z from 0
 * to 1, poly vertices, the last equal to the first; the reference reads
 * NACA_filename with list-directed input).  It is shifted to the quarter
 * chord, rotated by -AOA degrees -- ay is rotated with the already rotated
 * az, so the shape is sheared by cos(AOA) as in the reference -- and
 * placed at (zc, yc) (variables_module.f90:144-146).  Yn, Zn: the node
 * arrays Y(-1:ny+3), Z(-1:nz+3) of the gridder.  Each (z,y) cell centre
 * gets the even-odd crossing parity, each cell holding a polygon vertex the
 * inside-fraction of its nsub x nsub sub-cells (nSubGrids = 50, at most
 * 256), and the plane is extruded along x into the interior of ETA.
 * Computed on the device into the resident ETA; with ETA non-NULL (the
 * caller's halo array) the caller's array is uploaded first and copied back
 * after, so its ghosts keep their values.  A bound the reference's box
 * search (airfoil.f90:614-654) does not find is 0, the module variable's
 * first value, so a polygon wholly outside the grid gives ETA = 0 (the
 * unit cavity of the reference, airfoil at (4, 4)); a polygon straddling
 * the lower edge of the grid, for which the reference indexes outside its
 * arrays, is refused with NTUST_E_ARG. */
int ntust_ray_casting(const double *az0, const double *ay0, int poly, double AOA, double zc,
                      double yc, const double *Yn, const double *Zn, int nsub, double *ETA);

/* dynamic_AOA (airfoil.f90:311-321): AOA = AOA1 + 10 |sin(2 PI f (time -
 * dt (StartDynamic - 1)))| with f = 2/PI and PI the single-precision
 * literal 3.14159265359 widened, as the reference evaluates it.  Host
 * only; the caller then calls ntust_ray_casting (main.f90:144-155). */
int ntust_dynamic_aoa(double AOA1, double time, double dt, double StartDynamic, double *AOA);

/* ---- Lagrangian-marker direct forcing --------------------------------------
 *
 * No counterpart in the reference, whose immersed boundary is the Eulerian
 * volume fraction ETA blended into the velocity after the projection
 * (calcul_new_velocity.f90:61-100).  With markers set, the resident step
 * forces the post-projection velocity u1 at the same place instead:
 * U = I u1 at the markers, F = -U/dt (a resting body), u2 = u1 + dt S F,
 * and FY, FZ = S F feed the force integral (virtualForceIntegrator.f90:22-81)
 * as before.  Keep ETA = 0 with markers (the two forcings are alternatives).
 * Markers lie on a curve in the (z, y) plane and are extruded along x: one
 * marker (m, ix) at (xc(ix), ym[m], zm[m]) for every x cell ix, of weight
 * wm[m] * iDx(ix).  I and S use the cubic B-spline kernel (support 2
 * widths, twice differentiable) with widths hx, hy, hz per axis, on each component's
 * staggered points, and are transposes in the volume-weighted products;
 * all sums run in a fixed order (reproducible bit for bit).  Xn, Yn, Zn:
 * the node arrays X(-1:nx+3), Y(-1:ny+3), Z(-1:nz+3) of the gridder.  A
 * grid finer than a kernel width at a marker (more than 4 points within
 * 2 h) is refused (NTUST_E_ARG).  n = 0 removes the markers. */
int ntust_set_markers(int n, const double *zm, const double *ym, const double *wm, double hx,
                      double hy, double hz, const double *Xn, const double *Yn, const double *Zn);

/* The marker velocities U of the last step, U[(a*n + m)*nx + ix-1] for
 * component a (0 u, 1 v, 2 w). */
int ntust_marker_velocity(double *U);

/* ---- discrete adjoint of the resident step ------------------------------
 *
 * No counterpart in the reference: gradients of an objective of the flow
 * with respect to the state and to a body force f (a per-face force field,
 * the control) that the resident step adds to the predictor,
 * u* = u* + dt*f on the interior (0 unless set).  The adjoint of a step is
 * the exact transpose of the step the forward ran: the transposed QUICK /
 * upwind predictor with the upwind choices frozen, the linearised
 * Smagorinsky viscosity, the SOR sweeps transposed one by one (same sweep
 * count; the operator is symmetric, so each transposed sweep is a
 * homogeneous SOR sweep with the colours in reverse order), the projection
 * with the ETA blend, and the boundary conditions.  Needs the red-black or
 * mixed pressure solver.  Adjoint arrays are halo arrays like the state. */

/* Body force (three halo arrays, interior used); all NULL removes it. */
int ntust_set_body_force(const double *fu, const double *fv, const double *fw);

/* Kinetic energy 0.5*sum(u^2+v^2+w^2)*iDx*iDy*iDz over the interior faces of
 * the device state. */
int ntust_kinetic_energy(double *ke);

/* Exact FP64 copies of the device state (u, v, w, p) in device memory. */
int ntust_checkpoint_reserve(int nslots);
int ntust_checkpoint_save(int slot);
int ntust_checkpoint_load(int slot);

/* Adjoint state lambda = dJ/d(u, v, w, p) at the current step boundary.
 * ntust_adjoint_set uploads it (a NULL array is zero) and clears the
 * body-force gradient; ntust_adjoint_get downloads it (NULL skips). */
int ntust_adjoint_set(const double *lu, const double *lv, const double *lw, const double *lp);
int ntust_adjoint_get(double *lu, double *lv, double *lw, double *lp);

/* lambda_u += scale * dKE/du (v, w alike) at the current device state. */
int ntust_adjoint_add_ke(double scale);

/* One reverse step.  The device state must be the state at the start of the
 * step; the step is recomputed (its sweep count in *ik), lambda becomes the
 * adjoint of that state, dt*lambda_u* is added to the body-force gradient,
 * and the device state is left unchanged. */
int ntust_adjoint_step(int *ik);

/* Body-force gradient dJ/df accumulated since the last ntust_adjoint_set or
 * ntust_adjoint_run (halo arrays, interior meaningful; NULL skips). */
int ntust_adjoint_gradient(double *gfu, double *gfv, double *gfw);

/* Objective of ntust_adjoint_run (no counterpart in the reference):
 *   J = mean over the last navg steps of
 *       w_ke * KE + w_cd * C_D + w_cl * C_L
 * with KE the kinetic energy after the step (ntust_kinetic_energy) and
 * C_D, C_L the step's drag and lift from the force integral
 * (virtualForceIntegrator.f90:22-81, the drag from FZ, the lift from FY,
 * as ntust_step reports them; ntust_configure's lx).  The force integral is
 * linear in the forcing once its skip tests (a line sum skips a sample that
 * is exactly zero) are taken as they fell in the step, and its adjoint uses
 * them so.  eta_gradient != 0: the reverse pass also accumulates dJ/dETA
 * (ntust_adjoint_eta_gradient), the sensitivity to the immersed-boundary
 * volume fraction through the direct forcing and the force integral; with
 * the ETA of a geometry at two angles of attack it gives dJ/dAOA by the
 * chain rule.  Default: (1, 0, 0, 0), the kinetic energy alone.  Kept
 * until changed or ntust_finalize. */
int ntust_adjoint_objective(double w_ke, double w_cd, double w_cl, int eta_gradient);

/* With markers and eta_gradient set: dJ/dZ_m and dJ/dY_m, the sensitivity
 * of J to the position of marker m in the (z, y) plane (its x copies
 * summed), accumulated like dJ/dETA; with the markers of a geometry and
 * their motion dX_m/dAOA it gives dJ/dAOA by the chain rule.  No
 * counterpart in the reference. */
int ntust_adjoint_marker_gradient(double *gz, double *gy);

/* dJ/dETA accumulated by the last ntust_adjoint_run (or by the
 * ntust_adjoint_step calls since ntust_adjoint_set), a halo array with the
 * interior meaningful.  NTUST_E_STATE unless ntust_adjoint_objective asked
 * for it. */
int ntust_adjoint_eta_gradient(double *geta);

/* Forward nsteps from the device state and back.  Objective: J of
 * ntust_adjoint_objective over the last navg steps.  The forward pass
 * keeps a checkpoint every checkpoint_every steps; the reverse pass
 * recomputes each segment from its checkpoint (the last segment's states
 * are kept from the forward pass) and applies the reverse steps.  Outputs:
 * *J, ik[nsteps] the sweep counts (may be NULL), ms[4] (may be NULL): forward
 * pass, reverse pass, its recomputation part, its adjoint part, in
 * milliseconds.  Afterwards the device holds the final forward state,
 * lambda the adjoint of the initial state and ntust_adjoint_gradient dJ/df.
 * Returns NTUST_E_STATE if a recomputed step's sweep count differs.  With
 * NTUST_PROGRESS set in the environment it writes one line per segment of
 * each pass to stderr. */
int ntust_adjoint_run(int nsteps, int checkpoint_every, int navg, double *J, int *ik,
                      double *ms);

/* Checkpoint spacing for ntust_adjoint_run (no counterpart in the
 * reference), for the current grid, objective and ETA-gradient setting:
 * *every = nsteps -- every step's state kept from the forward pass, no
 * recomputation -- when the run's stores for that spacing fit in the
 * device's free memory (plus what those stores already hold) with
 * reserve_bytes left over; else ceil(sqrt(nsteps)).  *bytes (may be NULL)
 * gets the stores' size for the chosen spacing.  Allocates nothing; the
 * run's other buffers (the step's and the adjoint's working arrays) are the
 * caller's to cover with reserve_bytes. */
int ntust_adjoint_checkpoint_plan(int nsteps, double reserve_bytes, int *every, double *bytes);

/* ---- measurement -------------------------------------------------------- */

/* The library's HIP stream (hipStream_t), for event timing by the caller. */
void *ntust_stream(void);

/* Timing of the dominant kernel, the SOR half-sweep: batches of sweeps in
 * which every sweep ran are bracketed with HIP events on the library
 * stream.  ntust_prof_reset zeroes the counters; ntust_prof_read returns the
 * summed milliseconds and the number of half-sweep launches they cover. */
int ntust_prof_reset(void);
int ntust_prof_read(double *ms, long long *launches);

/* The timed kernel: returns 0 two-pass FP64 half-sweep, 2 two-pass FP32
 * half-sweep, 6 the lexicographic step; *bytes_per_cell gets its compulsory
 * bytes per grid cell and launch, averaged over a sweep (FP64: read p and b,
 * write p; FP32: read and write e, read r on the colour-0 cells, the
 * colour-1 sweeps being shifted; lexicographic: 24 B per cell update, over
 * the timed launches). */
int ntust_prof_kernel(double *bytes_per_cell);

/* Arithmetic of the resident step's QUICK predictor and Smagorinsky
 * viscosity (NTUST_ARITH_*; no counterpart in the reference, whose order
 * NTUST_ARITH_EXACT keeps).  The fast modes evaluate the same
 * discretisation (discretisation_QUICK_centre.f90:18-112,
 * smagorinsky.f90:14-24,60-79) with precomputed reciprocal tables and
 * fused multiply-adds; the upwind predictor and the drop-in entry points
 * always run exact.  Kept until ntust_finalize. */
int ntust_set_arithmetic(int mode);

/* Measurement aid (no counterpart in the reference): copies nbytes (rounded
 * down to 16-byte words, 16 B .. 64 GB) between two fresh device buffers
 * with a plain streaming kernel, k_stream_copy, each byte read once and
 * written once, and returns its time (*ms may be NULL).  A counter profile
 * of the same process then holds a copy of known traffic to check the
 * memory counters against.  Allocates and frees the two buffers. */
int ntust_prof_copy(double nbytes, double *ms);


#ifdef __cplusplus
}
#endif
#endif
