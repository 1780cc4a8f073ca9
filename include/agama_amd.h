/*  agama_amd: MI355X (gfx950) Multipole potential evaluation and DOP853 orbit integration.

    Plain C ABI: opaque handles, plain pointers and sizes, integer status codes, a
    thread-local last-error string.  Every entry point names the reference (Agama) interface
    that it replaces; paths are relative to the reference source tree.

    Conventions (as in the reference):
      - positions/velocities are Cartesian, rows of 3 (x,y,z) or 6 (x,y,z,vx,vy,vz), row-major;
      - "grad" outputs are +dPhi/dx (the reference's BasePotential::eval gradient), while the
        NEMO-style entry point returns accelerations -dPhi/dx;
      - a non-zero Omega integrates in a frame rotating about z, inputs and outputs are given
        in the rotating frame (src/orbit.h:219-232).
    All computation runs on the current HIP device of the calling thread at creation time.
    Functions taking host pointers copy to and from the device and synchronise; the _device
    variants take device pointers and a hipStream_t (passed as void*) and are asynchronous.
*/
#ifndef AGAMA_AMD_H
#define AGAMA_AMD_H
#ifndef __cplusplus
#include <stdbool.h>
#endif
#ifdef __cplusplus
extern "C" {
#endif

typedef struct agx_potential agx_potential;

/* status codes */
#define AGX_OK            0
#define AGX_EINVAL        1   /* invalid argument */
#define AGX_EDEVICE       2   /* HIP runtime error */
#define AGX_EUNSUPPORTED  3   /* valid in the reference, not provided by this library */

/* flags: precision of the Multipole evaluation ... */
#define AGX_PRECISION_FP64 0
#define AGX_PRECISION_FP32 1  /* splines, harmonic scaling and Fourier synthesis in float from float
                                 tables; coordinates, extrapolation, stages, error control and
                                 energy stay FP64 */
#define AGX_PRECISION_MASK 0xF
/* ... and, for the orbit entry points, the processing order: by default the library orders
   the batch longest-expected-integration first (|T| / dynamical time at the start point);
   AGX_ORDER_AS_GIVEN keeps the input order (or the caller's d_order). */
#define AGX_ORDER_AS_GIVEN 0x100
/* ... and the integration method (OrbitIntParams::Method, src/orbit.h:116-129): DOP853 by
   default, or the 8th-order Runge-Kutta-Nystrom method DPRKN8 (src/math_ode.cpp:426-701),
   which evaluates positions only (9 evaluations per step attempt instead of 12).  The DPRKN8
   kernel gathers FP64 node records with 16-byte loads whatever agx_set_fetch chose. */
#define AGX_METHOD_DPRKN8 0x200

/* Message of the last failed call on this thread ("" if none).
   Replaces agama_getError (src/interface_c.cpp:15-19), which is a process-global string. */
const char* agx_last_error(void);

/* Create a potential from key=value parameters, separated by commas or newlines
   ("type=Multipole, file=coefs.ini" or "type=Multipole, density=Dehnen, gamma=1, ...").
   Replaces agama_createPotential (src/interface_c.h:46, src/interface_c.cpp:31-54) and the
   Multipole part of potential::createPotential (src/potential_factory.cpp:1357-1441).
   A density is expanded by the Poisson solver (computePotentialCoefsFromSource,
   src/potential_multipole.cpp:1013-1114) on the device: density at the quadrature points,
   forward spherical-harmonic transform and radial integrals (agx_set_construction).
   Returns NULL on error. */
agx_potential* agx_potential_create(const char* params);

/* Create a Multipole potential directly from coefficient arrays: radii[nr] and
   phi/dphi[c*nr + k] for c = l*(l+1)+m < (lmax+1)^2
   (reference: Multipole::Multipole, src/potential_multipole.cpp:1491-1516). */
agx_potential* agx_potential_from_coefs(int nr, int lmax, const double* radii,
    const double* phi, const double* dphi);

/* Replaces agama_deletePotential (src/interface_c.cpp:56-60). */
void agx_potential_destroy(agx_potential* pot);

/* Build a Multipole expansion from key=value parameters as agx_potential_create does and
   write its coefficient file: the reference's Potential(type='Multipole', density=...)
   .export(path) sequence (src/potential_multipole.cpp:1400-1441,
   src/potential_factory.cpp:909-929).  Needs no GPU unless agx_set_construction chose the
   device. */
int agx_build_coef_file(const char* params, const char* path);

/* Where the Poisson solver of a density runs: AGX_CONSTRUCT_DEVICE, AGX_CONSTRUCT_HOST (on the
   CPU in the reference's operation order), AGX_CONSTRUCT_HOST_STAGED (on the CPU in the
   device's order, the check of the device kernels), or AGX_CONSTRUCT_DEFAULT: the device for
   agx_potential_create, the host for agx_build_coef_file.  Returns AGX_OK, or AGX_EINVAL for
   another value. */
#define AGX_CONSTRUCT_DEFAULT    -1
#define AGX_CONSTRUCT_DEVICE      0
#define AGX_CONSTRUCT_HOST        1
#define AGX_CONSTRUCT_HOST_STAGED 2
int agx_set_construction(int mode);

/* Host-only diagnostic: the interpolation tables built from a coefficient file, in the
   reference's per-harmonic layout.  which: 0 = ln r grid, 1 = tau grid, 2 = the 9 node arrays
   (f, f_x, f_y, f_xx, f_xy, f_yy, f_xxy, f_xyy, f_xxyy, each nr*nt row-major) of harmonic m,
   3 = {logScaling, invPhi0, lmax, mmax, mmin, step}, 4 = the 1-d variant's f, f_x, f_xx of
   term c = m, 5 = the per-cell polynomial form of harmonic m used by the FP64 kernels (36
   coefficients c[a][b] of u^a v^b per cell, cells (i, j) row-major over (nr-1) x (nt-1)).
   Returns the number of doubles written (or needed when out is NULL), or -1. */
long agx_debug_tables(const char* coef_path, int which, int m, double* out);

/* Write the coefficients in the reference's text format
   (writePotential / writePotentialMultipole, src/potential_factory.cpp:909-929). */
int agx_potential_write(const agx_potential* pot, const char* path);

/* Query the coefficient arrays: returns nr and lmax; copies arrays when pointers are non-NULL
   (sizes nr and (lmax+1)^2*nr).  Replaces Multipole::getCoefs' data access. */
int agx_potential_coefs(const agx_potential* pot, int* nr, int* lmax, double* radii,
    double* phi, double* dphi);

/* Single-point potential, gradient (+dPhi/dx) and, when deriv2 is non-NULL, Hessian
   {dx2, dy2, dz2, dxdy, dydz, dxdz} of Phi.
   Replaces agama_evalPotential (src/interface_c.h:75, src/interface_c.cpp:68-98). */
double agx_eval_potential(const agx_potential* pot, const double pos[3], double t,
    double deriv[3], double deriv2[6]);

/* Batched potential and gradient at n points (host buffers pos[n*3], phi[n], grad[n*3];
   phi or grad may be NULL).  Replaces Potential.eval/potential/force over
   BatchFunction (src/interface_python.cpp:3027-3207, 1206-1489). */
int agx_eval_many(const agx_potential* pot, long n, const double* pos, double t,
    double* phi, double* grad, int flags);

/* Same with device buffers, asynchronous on `stream` (hipStream_t). */
int agx_eval_many_device(const agx_potential* pot, long n, const double* d_pos, double t,
    double* d_phi, double* d_grad, int flags, void* stream);

/* Batched potential, gradient and Hessian hess[n*6] = {dx2, dy2, dz2, dxdy, dydz, dxdz} of Phi
   (any output may be NULL; FP64 only: AGX_PRECISION_FP32 returns AGX_EUNSUPPORTED).  Replaces
   BasePotential::eval with deriv2 behind Potential.eval(..., der=True) and forceDeriv
   (src/interface_python.cpp:3027-3207, 3106-3112; the Python layer returns -hess, the force
   derivatives). */
int agx_eval_hess_many(const agx_potential* pot, long n, const double* pos, double t,
    double* phi, double* grad, double* hess, int flags);
int agx_eval_hess_many_device(const agx_potential* pot, long n, const double* d_pos, double t,
    double* d_phi, double* d_grad, double* d_hess, int flags, void* stream);

/* Batch orbit integration with DOP853 (or DPRKN8, flags) in the given potential, restating agama.orbit
   (src/interface_python.cpp:7689-8281) with dtype=float64 and separateTime output:
     ic[n*6], tstart[n] (NULL: 0), ttotal[n], trajsize[n] (>=1) on the host;
     out_time[rows], out_traj[rows*6], rows = sum(trajsize), orbit i at row offset
     sum_{j<i} trajsize[j], NaN-padded after an early stop (storeOrbitArrays,
     src/interface_python.cpp:6921-6959); out_final[n*6] = last recorded point;
     out_nsteps / out_nattempts / out_nevals: accepted steps, attempts, RHS evaluations;
     out_maxde: max |E/E0-1| over accepted steps (E - Omega*Lz when Omega != 0);
     out_status: 0 done, 1 stepper failure (orbit.cpp:69-75), 2 maxNumSteps reached.
   Any output pointer may be NULL.  Returns AGX_OK or an error code. */
int agx_orbit_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps,
    double* out_time, double* out_traj, double* out_final, long* out_nsteps,
    long* out_nattempts, long* out_nevals, double* out_maxde, int* out_status, int flags);

/* Same with device buffers.  d_traj_offset[n] gives each orbit's first output row; d_order
   (may be NULL) a processing order (permutation of 0..n-1); without it the library computes
   the longest-first order itself (stream-ordered workspace) unless flags has
   AGX_ORDER_AS_GIVEN.  Asynchronous on `stream`.  Calls that overlap on one potential handle race on its
   work-queue counter; use one handle per concurrent stream. */
int agx_orbit_batch_device(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
    const long* d_traj_offset, const int* d_order, double accuracy, double Omega,
    long maxNumSteps, double* d_time, double* d_traj, double* d_final, long* d_nsteps,
    long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags,
    void* stream);

/* agx_orbit_batch_device without any allocation: no counterpart in the reference; for callers
   that cannot allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  Same arguments, results and stream semantics as
   agx_orbit_batch_device (the same kernels, bit for bit); the longest-first schedule's workspace
   (sort keys, radix-sort storage, work order) lives in a static device array of 48 MiB that the
   library holds on each device it runs on until the process ends.  At most 1,048,576 orbits
   per call (AGX_EINVAL beyond); the potential's device must be current; plain DOP853/DPRKN8
   integration only (no deviation vectors, targets or every-step recording).  Calls that overlap
   on one device race on that array (and, as for every orbit entry point, on the potential's
   work-queue counter). */
int agx_orbit_batch_device_noalloc(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
    const long* d_traj_offset, const int* d_order, double accuracy, double Omega,
    long maxNumSteps, double* d_time, double* d_traj, double* d_final, long* d_nsteps,
    long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags,
    void* stream);

/* agx_orbit_batch with the variational equation (src/orbit_variational.cpp RuntimeVariational,
   deviation vectors advanced with Ode2StepperGL4<3>, src/math_ode.cpp:1017-1250):
   out_dev (agama.orbit(der=True)): the six deviation vectors started from the identity, sampled
     at the trajectory's times: per orbit at row offset sum_{j<i} trajsize[j] * 36,
     [vector][row][6] in the rotating frame, NaN-padded like the trajectory; vector k is the
     derivative of the phase-space point with respect to initial coordinate k;
   out_lyap[n*2] (agama.orbit(lyapunov=True)): the Lyapunov exponent in units of the orbital
     period T_circ(E) of the starting point and the onset time of chaos in the same unit
     (calcLyapunovExponent, orbit_variational.cpp:360-448): 0 and infinity for regular orbits,
     NaN for orbits with fewer than 5 bins of the deviation-vector growth.  Without out_dev the
     growth is followed with one vector, as the reference does.  The fit is a Levenberg-Marquardt
     of this library (the reference uses GSL's); orbits that stop early are integrated twice.
   At least one of out_dev, out_lyap must be non-NULL.  DOP853 and FP64 only (other flags return
   AGX_EUNSUPPORTED). */
int agx_orbit_var_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps,
    double* out_time, double* out_traj, double* out_final, double* out_dev, double* out_lyap,
    long* out_nsteps, long* out_nattempts, long* out_nevals, double* out_maxde, int* out_status, int flags);
/* The fitting step of the Lyapunov exponent on already binned data (host only): t[nbins],
   lnw[nbins] the bin times and log mean deviation magnitudes, max_time the last recorded time,
   period T_circ; out[2] as out_lyap above.  Replaces the fit and decisions of
   calcLyapunovExponent (src/orbit_variational.cpp:406-448). */
int agx_lyapunov_fit(const double* t, const double* lnw, int nbins, double max_time, double period,
    double* out);
/* Deviation vectors (out_dev required) with device buffers, asynchronous on `stream`; the
   per-lane state (1.3 KB per resident lane) is allocated stream-ordered.  The Lyapunov fit
   needs the host form. */
int agx_orbit_var_batch_device(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
    const long* d_traj_offset, const int* d_order, double accuracy, double Omega,
    long maxNumSteps, double* d_time, double* d_traj, double* d_final, double* d_dev, long* d_nsteps,
    long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags, void* stream);

/* ---- every-step trajectories and interpolated orbits (agama.orbit with trajsize=0 or
   dtype=object; src/interface_python.cpp:7577-7584, 7864-7869, 8121-8127, 6605-6880) ---- */
typedef struct agx_orbit_record agx_orbit_record;
/* Integrate a batch and keep what it recorded in device memory: orbit i records trajsize[i]
   rows at regular intervals (as agx_orbit_batch) or, when trajsize[i] == 0, the start and the
   end of every accepted step (RuntimeTrajectory with sampling interval 0, src/orbit.cpp:38-45),
   in the rotating frame; der != 0 also records the six deviation vectors at the same times
   (DOP853, FP64 only, as agx_orbit_var_batch).  out_rows[n]: rows of orbit i (trajsize[i], or
   the recorded count when it is 0).  The other outputs are those of agx_orbit_batch; any may be
   NULL.  Returns NULL on error (agx_last_error).  Every-step orbits share a budget of 2^28
   rows (2^25 with der): a block per orbit, and an overflow pool of chunks that a lane takes on
   the device when its orbit outgrows the block; an orbit that finds the pool exhausted is
   integrated a second time with a block of its size.  The rows are then gathered into one
   compact buffer. */
agx_orbit_record* agx_orbit_record_batch(const agx_potential* pot, long n, const double* ic,
    const double* tstart, const double* ttotal, const int* trajsize, double accuracy, double Omega,
    long maxNumSteps, int der, long* out_rows, double* out_final, long* out_nsteps, long* out_nattempts,
    long* out_nevals, double* out_maxde, int* out_status, int flags);
/* The recorded rows, orbit after orbit (orbit i at row offset sum_{j<i} out_rows[j]):
   out_time[rows], out_traj[rows*6], out_dev[rows*36] as [vector][row][6] per orbit (any may
   be NULL; out_dev needs der).  The rows are gathered on the device and copied once. */
int agx_orbit_record_fetch(const agx_orbit_record* rec, double* out_time, double* out_traj,
    double* out_dev);
/* Interpolated orbits (agama.Orbit, createOrbitObject, interface_python.cpp:6793-6880): each
   series - the trajectory of orbit i (series i) or its k-th deviation vector (series n + 6*i + k,
   with der) - becomes three quintic splines x(t), y(t), z(t) in the inertial frame through the
   recorded positions with the velocities as derivatives (repeated timestamps skipped, time
   order restored for backward integration); built on the device on first use.
   num_series: n or 7n.  series_info: nodes per series (the Orbit's length) and whether it was
   recorded backward in time.  node_times: the node times of one series in increasing order. */
long agx_orbit_record_num_series(const agx_orbit_record* rec);
/* Rows the first pass of agx_orbit_record_batch shares among its every-step orbits (0: the
   default 2^28, or 2^25 with deviation vectors); a small budget sends orbits through the
   overflow pool and, when that runs out, through the second integration. */
int agx_set_record_rows(long rows);
int agx_orbit_record_series_info(agx_orbit_record* rec, long* out_nodes, int* out_reversed);
int agx_orbit_record_node_times(agx_orbit_record* rec, long series, double* out_t);
/* The spline nodes of one series (no counterpart in the reference's C interface; the data of
   its QuinticSpline x(t), y(t), z(t), src/math_spline.cpp:1391-1429): out[nodes*10], per node
   t, x, y, z, dx/dt, dy/dt, dz/dt, d2x/dt2, d2y/dt2, d2z/dt2 in the inertial frame, in order of
   increasing t (nodes from agx_orbit_record_series_info). */
int agx_orbit_record_nodes(agx_orbit_record* rec, long series, double* out);
/* Evaluate m queries (series[q], t[q]) on the device.  mode 0 (Orbit.__call__,
   interface_python.cpp:6643-6714): out[m*6] position and velocity rotated back into the
   rotating frame at angle Omega*t, NaN outside the recorded time span; mode 1 (the x, y, z
   Spline members): out[m*9] = x, y, z, their first and their second derivatives in the inertial
   frame, continued linearly beyond the ends (QuinticSpline::evalDeriv, math_spline.cpp:1405-1429). */
int agx_orbit_record_interp(agx_orbit_record* rec, long m, const long* series, const double* t, int mode,
    double* out);
void agx_orbit_record_free(agx_orbit_record* rec);

/* ---- target plug-ins (agama.Target, src/interface_python.cpp:7005-7250, and orbit(targets=),
   :8115-8120, 8153-8162) ---- */
typedef struct agx_target agx_target;
/* TargetDensitySphHarm (src/galaxymodel_densitygrid.cpp:356-411): density discretised as
   linear B-splines on the radial grid (a grid starting at 0 loses that node; the origin is
   always a node of the l=0 term) times the spherical harmonics of even l <= lmax and even
   0 <= m <= mmax.  NULL on invalid parameters. */
agx_target* agx_target_density_sphharm(int lmax, int mmax, const double* gridr, int ngridr);
/* TargetDensityCylindrical<degree> (src/galaxymodel_densitygrid.cpp:486-545): degree 0
   (DensityCylindricalTopHat) constant cells, degree 1 (DensityCylindricalLinear) bilinear
   B-splines on the (R, |z|) grid (grids starting at 0 lose that node; R = 0 and z = 0 are
   implied), times the azimuthal Fourier terms cos(m phi) of even 0 <= m <= mmax.  NULL on
   invalid parameters. */
agx_target* agx_target_density_cylindrical(int degree, int mmax, const double* gridR, int ngridR,
    const double* gridz, int ngridz);
/* TargetDensityClassic<degree> (DensityClassicTopHat = 0 / DensityClassicLinear = 1,
   src/galaxymodel_densitygrid.cpp:146-239): spheroidal shells on gridr (a leading 0 is dropped)
   with axis ratios y/x and z/x, each shell cut into three panes of stripsPerPane^2 cells; top-hat
   cells or trilinear nodes.  NULL on invalid parameters. */
agx_target* agx_target_density_classic(int degree, int stripsPerPane, const double* gridr, int ngridr,
    double axisRatioY, double axisRatioZ);
/* TargetKinemShell<degree> (src/galaxymodel_losvd.h:121-136, .cpp:568-576): the radial and
   tangential velocity squared, v_r^2 and v_t^2, each on the B-spline basis of degree 0..3 over
   gridr (from r = 0), 2 * (ngridr + degree - 1) values.  NULL on invalid parameters. */
agx_target* agx_target_kinem_shell(int degree, const double* gridr, int ngridr);
/* number of values a target records per orbit (numCoefs), -1 for NULL */
int agx_target_num_values(const agx_target* t);
void agx_target_destroy(agx_target* t);
/* Integrate a batch (DOP853, FP64) feeding 1..4 targets: per accepted step ten equally spaced
   dense-output points, each weighted by a tenth of the step (RuntimeFncTarget,
   src/galaxymodel_target.h:95-155), accumulated per orbit in FP64 on the device;
   out_targets[t][n * numValues] receives the datacube in float divided by the orbit's
   integration time (unchanged when it is 0).  out_final/out_nsteps/out_status as in
   agx_orbit_batch (may be NULL).  flags: AGX_METHOD_DPRKN8 selects DPRKN8 (its dense output
   feeds the samples).  FP32 tables and deviation vectors are not combined with targets
   (AGX_EUNSUPPORTED). */
int agx_orbit_target_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, double accuracy, double Omega, long maxNumSteps, int ntargets,
    const agx_target* const* targets, float* const* out_targets, double* out_final, long* out_nsteps,
    int* out_status, int flags);
/* The same with trajectories recorded in the same integration (interface_python.cpp:7991-7994:
   the target matrices, then the trajectory outputs): trajsize[n] >= 1 rows per orbit at regular
   intervals, stored as in agx_orbit_batch (out_time[rows], out_traj[rows * 6], orbit after orbit;
   either may be NULL). */
int agx_orbit_target_traj_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps, int ntargets,
    const agx_target* const* targets, float* const* out_targets, double* out_time, double* out_traj,
    double* out_final, long* out_nsteps, int* out_status, int flags);

/* Launch geometry of the orbit kernel: number of workgroups (0 = one full wave of resident
   workgroups on every CU) and threads per workgroup (0 or 256, the only size it is built for).
   Set by the autotuner (agama_amd.tune). */
int agx_set_orbit_launch(int num_blocks, int block_size);

/* How the FP64 kernels read the Multipole's 2-D splines, for the evaluation kernel and for the
   orbit kernel: from the node records, 0 = one 8-byte load per value and lane, 1 = 16-byte
   loads per lane, 2 = a wavefront-cooperative gather through LDS; 3 = from the per-cell
   polynomial form of the same splines.  Modes 0-2 give identical results; mode 3 agrees with
   them to rounding (it also takes the azimuth's cos/sin from x/R, y/R, uses reciprocals instead
   of divisions and a tabulated Legendre recurrence in the power-law regions).  Default: 3, 3
   (the autotuner, agama_amd.tune, may choose another). */
int agx_set_fetch(int eval_mode, int orbit_mode);

/* ---- NEMO acceleration plugin (src/interface_nemo.cpp:47-60, 114-220) ----
   iniacceleration(pars, npar, file, &accel, &need_m, &need_v) (replaces interface_nemo.cpp:
   209-220): pars[0] = Omega (pattern speed), pars[1] = G (only 1: other values stop the process
   with a message, as a missing file does); file names a Multipole coefficient file.  The
   returned accel has the acc_pter signature:
     accel(ndim, t, n, m, pos, vel, flags, pot, acc, flag, numtype)
   with ndim = 2 or 3, numtype 'f' or 'd', flag bit 0: add to pot, bit 1: add to acc (else assign),
   flags[i] & 1 selects active bodies (NULL: all); positions and accelerations are in the inertial
   frame, the potential rotating at Omega (its frame at angle Omega*t).  The active bodies of a
   call are evaluated as one batch on the device.  */
typedef void (*agx_acc_pter)(int ndim, double t, int n, const void* m, const void* pos,
    const void* vel, const int* flags, void* pot, void* acc, int flag, char numtype);
void iniacceleration(const double* pars, int npar, const char* file, agx_acc_pter* accel,
    bool* need_m, bool* need_v);
/* the older interface (replaces interface_nemo.cpp:180-206): installs the potential as above */
void inipotential(const int* npar, const double* pars, const char* file);
/* potential and acceleration of one body, assigned (replace interface_nemo.cpp:150-177) */
void potential_double(const int* ndim, const double* pos, double* acc, double* pot, const double* time);
void potential_float(const int* ndim, const float* pos, float* acc, float* pot, const float* time);

/* ---- drop-in aliases of the reference's plain C interface (src/interface_c.h:15-85) ---- */
typedef agx_potential agama_Potential;
const char* agama_getError(void);
agama_Potential* agama_createPotential(const char* params);
void agama_deletePotential(agama_Potential* pot);
double agama_evalPotential(const agama_Potential* pot, const double pos[3], double time,
    double deriv[3], double deriv2[6]);



/* Orbit-kernel launch configuration chosen by the autotuner (agama_amd/tune.py): wavefronts
   per SIMD the persistent kernel is compiled for (1: 512 registers per lane, 2: 256), and the
   current settings of agx_set_orbit_launch / agx_set_fetch / agx_set_orbit_waves. */
int agx_set_orbit_waves(int waves);
/* Group mode of the DOP853 kernel (no counterpart in the reference: a scheduling knob of this
   library).  In batches of at least min_batch orbits (FP64; DOP853 on node records or per-cell
   tables (orbit fetch mode 1 or 3) at 1 wave/SIMD, or DPRKN8 on per-cell tables; no NFW
   halo, no targets, deviation vectors or every-step recording), one wavefront in `stride`
   integrates the first `orbits` orbits of the processing order - the longest expected - with 8
   lanes per orbit, each lane evaluating a share of the azimuthal harmonics; orbits = -1: one per
   8 lanes of those wavefronts; stride = 0: off; stride = -1: automatic (every orbit in groups when
   they fit, 4 for up to 2 orbits per lane, 8 for up to 4, above that 8 after a work-order probe
   (agx_set_orbit_probe) and off without one; after a probe, with the default `orbits`, the
   group wavefronts' orbits are dealt round-robin so that each holds one of the longest).  Results
   differ from the one-lane form only by the order of the additions in the Fourier sum and in the
   power-law extrapolation's sum over harmonics.  Defaults: -1, -1, 65536. */
int agx_set_orbit_groups(int stride, long orbits, long min_batch);
/* Work-order probe of the DOP853/DPRKN8 batches (no counterpart in the reference: a scheduling
   knob of this library).  In batches of at least min_batch orbits (without deviation vectors,
   targets or every-step recording) every orbit is first integrated for at most `steps` steps
   with no output but the time it reached, and the orbits are processed in the order of the
   steps they would take at that rate, longest first (otherwise: |T| / t_dyn at the start);
   with the probe, group mode's automatic stride above 4 orbits per lane is 8.  Results do not
   depend on the order except through group mode.  steps = 0: off.  Defaults: 64, 65536. */
int agx_set_orbit_probe(int steps, long min_batch);
/* The current settings of agx_set_orbit_groups and agx_set_orbit_probe (any pointer may be
   NULL; no counterpart in the reference), for the autotuner to restore them. */
int agx_get_orbit_schedule(int* group_stride, long* group_orbits, long* group_min_batch, int* probe_steps,
    long* probe_min_batch);
int agx_get_orbit_config(int* num_blocks, int* fetch_eval, int* fetch_orbit, int* waves);

/* One integration of a fixed batch of device arrays (arguments as agx_orbit_batch_device)
   captured as a HIP graph - a queue-counter fill kernel and the persistent kernel - for
   callers that integrate the same batch layout repeatedly (e.g. after updating d_ic in
   place).  The work order is computed once, at creation, from the initial conditions then
   present; the graph owns its queue counter, so graphs of one potential may run
   concurrently.  No counterpart in the reference (its orbit() is host-synchronous). */
typedef struct agx_orbit_graph agx_orbit_graph;
agx_orbit_graph* agx_orbit_graph_create(const agx_potential* pot, long n, const double* d_ic,
                                        const double* d_tstart, const double* d_ttotal,
                                        const int* d_trajsize, const long* d_traj_offset,
                                        double accuracy, double Omega, long maxNumSteps,
                                        double* d_time, double* d_traj, double* d_final, long* d_nsteps,
                                        long* d_nattempts, long* d_nevals, double* d_maxde,
                                        int* d_status, int flags);
int agx_orbit_graph_launch(agx_orbit_graph* g, void* stream);
void agx_orbit_graph_destroy(agx_orbit_graph* g);

/* ---- MLP potential surrogate (no counterpart in the reference) ----
   Phi_NN(x) = Phi_base(x) (1 + out_scale * MLP(|x|/r, |y|/r, |z|/r, ln r / ln_scale)), with
   Phi_base the potential's expansion truncated to l <= base_lmax (FP64) and MLP a 4-256-256-
   256-256-1 SiLU network evaluated in bf16 on MFMA with fp32 accumulation; the force is the
   exact gradient of the network by a hand-written backward.  Parameters: AGX_MLP_NPARAM
   floats, W1[256x4] b1 W2[256x256] b2 W3 b3 W4 b4 w5[256] b5, weights row-major [out][in]. */
#define AGX_MLP_NPARAM_C 198913
typedef struct agx_surrogate agx_surrogate;
/* train on the GPU: params "samples=20e6, batch=16384, steps=6000, lr=1e-3, seed=1,
   baseLmax=2, rmin=1e-2, rmax=1e2, warmup=200" (all optional).  Samples have ln r uniform in
   [ln rmin, ln rmax] and isotropic directions; targets are the FP64 Multipole's relative
   residual over the base; Adam with cosine decay; one training step is a replayed HIP graph.
   baseLmax: 0 to 8, below the potential's own order (a higher base leaves a smaller residual
   and costs more FP64 work per evaluation).  The potential must outlive the surrogate. */
agx_surrogate* agx_surrogate_train(const agx_potential* pot, const char* params);
/* from given parameters (e.g. saved by a caller from agx_surrogate_info) */
agx_surrogate* agx_surrogate_create(const agx_potential* pot, const float* theta, int base_lmax,
                                    float ln_scale, float out_scale, double rmin, double rmax);
void agx_surrogate_destroy(agx_surrogate* s);
/* theta[AGX_MLP_NPARAM_C] and meta[8] = {base_lmax, ln_scale, out_scale, rmin, rmax,
   training steps, last training loss, AGX_MLP_NPARAM_C}; either may be NULL */
int agx_surrogate_info(const agx_surrogate* s, float* theta, double* meta);
/* training loss per step (mean squared error of the scaled residual); returns the count */
long agx_surrogate_loss_history(const agx_surrogate* s, float* out, long cap);
/* Phi_NN and +grad Phi_NN at n points (host / device arrays) */
int agx_surrogate_eval_many(const agx_surrogate* s, long n, const double* pos, double* phi, double* grad);
int agx_surrogate_eval_many_device(const agx_surrogate* s, long n, const double* d_pos,
                                   double* d_phi, double* d_grad, void* stream);
/* the raw network on features feat[n][4] (float): out[n] and d out / d feat [n][4] */
int agx_surrogate_network(const agx_surrogate* s, long n, const float* feat, float* out, float* dfeat);
/* one forward/backward of the training kernels on a batch (n a multiple of 512): mean
   squared error and its gradient w.r.t. the parameters (no update) */
int agx_surrogate_train_grad(const agx_surrogate* s, long n, const float* feat, const float* target,
                             float* loss, float* grad);
/* Orbit integration with the surrogate force, the arguments and outputs of agx_orbit_batch
   plus max_rel_de: every orbit is integrated with the surrogate while its energy error
   |E/E0 - 1|, measured with the FP64 Multipole at every accepted step, stays at or below
   max_rel_de and its evaluation points stay within the training radii; otherwise it is
   integrated again from its initial conditions with the FP64 Multipole (its status gets
   +8).  n_fallback (may be NULL) receives the number of such orbits.  The device variant
   synchronises the stream once, to size the fallback launch. */
int agx_surrogate_orbit_batch(const agx_surrogate* s, long n, const double* ic, const double* tstart,
                              const double* ttotal, const int* trajsize, double accuracy, double Omega,
                              long maxNumSteps, double max_rel_de, double* out_time, double* out_traj,
                              double* out_final, long* out_nsteps, long* out_nattempts, long* out_nevals,
                              double* out_maxde, int* out_status, long* n_fallback);
int agx_surrogate_orbit_batch_device(const agx_surrogate* s, long n, const double* d_ic,
                                     const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
                                     const long* d_traj_offset, double accuracy, double Omega,
                                     long maxNumSteps, double max_rel_de, double* d_time, double* d_traj,
                                     double* d_final, long* d_nsteps, long* d_nattempts, long* d_nevals,
                                     double* d_maxde, int* d_status, long* n_fallback, void* stream);
/* Batch routing of the two calls above (no counterpart in the reference, whose path has no
   surrogate): for n >= 8192 a pilot of every 64th orbit runs on the surrogate for at most 200
   steps first; when half of the pilot or more already falls back, every orbit of the batch is
   integrated with the FP64 Multipole directly (each counted and marked as a fallback) rather
   than twice.  The pilot launch ends as soon as half of its orbits have fallen back (the
   decision is then known).  enable = 0 turns the pilot off (default on); last_pilot reports the
   last batch's pilot size and the fallbacks counted until the pilot ended, on this thread (0, 0
   without a pilot). */
int agx_set_surrogate_pilot(int enable);
int agx_surrogate_last_pilot(long* out_pilot_orbits, long* out_pilot_fallbacks);
/* the training GEMM on host arrays: C[M][N] = sum_k A(m,k) B(k,n), bf16 bit patterns,
   A row-major [M][K] if a_kcontig else [K][M], B [N][K] if b_kcontig else [K][N] */
int agx_debug_gemm(int M, int N, int K, int splits, int a_kcontig, int b_kcontig,
                   const unsigned short* A, const unsigned short* B, float* C);

#ifdef __cplusplus
}
#endif
#endif
