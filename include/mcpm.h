/*
 * mcpm.h — C ABI of the MI355X particle Monte Carlo step (libmcpm.so).
 *
 * Plain C: POD structs, plain pointers and sizes, integer status codes.
 * Each entry point replaces one per-step operator of the reference program
 * mohsensadr/Monte-Carlo-Particle-Method; the comment above each names the
 * reference function (prototype in src/cpp_headers.h, body file:line).
 *
 * Ownership and residency.  The reference keeps particles in caller-owned
 * host arrays x2[], U1[], U2[], U3[] (main.cpp:104-165) and per-cell state
 * in struct CELLS.  Here a context owns the device copy: particles are
 * uploaded once, stay resident in HBM across steps and come back to the
 * host only on request.  Particles are kept physically ordered by cell
 * (a stable counting sort of the previous order), so the reference's
 * cells[c].indices_inside[u] is simply start[c] + u.
 *
 * Threading: one context is one HIP stream; calls on one context must not
 * overlap.  All calls return MCPM_OK or an error code; nothing exits the
 * process (the reference calls exit(1)/exit(2) on the same conditions).
 */
#ifndef MCPM_H
#define MCPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MCPM_OK            0
#define MCPM_ERR_ARG       1  /* bad argument or unsupported parameter combination */
#define MCPM_ERR_HIP       2  /* HIP runtime failure (no device, launch failure)    */
#define MCPM_ERR_CAPACITY  3  /* more particles / candidates than the context holds */
#define MCPM_ERR_DOMAIN    4  /* particle outside the domain: Main_Functions.cpp:2248-2263 */
#define MCPM_ERR_NAN       5  /* non-finite velocity: Fokker_Planck.cpp:1137-1149    */
#define MCPM_ERR_BC        6  /* >20 wall interactions in one step: Main_Functions.cpp:4958-4965 */
#define MCPM_ERR_STATE     7  /* call out of order (e.g. step before upload)       */

/* ---- enumerations ------------------------------------------------------ */
enum mcpm_precision { MCPM_FP64 = 0, MCPM_FP32 = 1 };
/* Boundary treatment along the slab axis (the reference's x2 / y axis).
   DIFFUSE: BC_thermal_wall_3D (Main_Functions.cpp:4917-5248), walls at
   wall_lo / wall_hi.  SPECULAR: BC_specular_reflection
   (Main_Functions.cpp:6399-6525).  PERIODIC: a periodic box, which the
   reference's 1D screened-Poisson path does not offer.  REINJECT (slab
   only): BC_evaporation of the vacuum problem (Main_Functions.cpp:5444-5617):
   a particle past wall_lo / wall_hi (the reference's 1e-15 and L - 1e-15)
   is put back at L/2 + U(-1,1) L/4 with a Maxwellian velocity at T_wall_lo
   (gas.T), and counted (mcpm_get_escaped).  The inverted problem's
   BC_inverted (Main_Functions.cpp:5251-5441) is DIFFUSE with walls at
   1e-15 and L - 1e-15 and T_wall_lo = gas.Tc, T_wall_hi = gas.Th. */
enum mcpm_boundary {
  MCPM_BC_DIFFUSE = 0,
  MCPM_BC_PERIODIC = 1,
  MCPM_BC_SPECULAR = 2,
  MCPM_BC_REINJECT = 3
};
/* Contact value Y(n): the reference's active polynomial
   (Main_Functions.cpp:1906-1915) or Carnahan-Starling (the commented
   alternative at Main_Functions.cpp:1912). */
enum mcpm_ymodel { MCPM_Y_REFERENCE = 0, MCPM_Y_CARNAHAN_STARLING = 1 };
/* Long-range attraction switch of velocity_update (Main_Functions.cpp:1612-1668):
   1 direct Sutherland integral (vlasov_taking_the_integral), 2 density
   expansion (KDE gradient), 3 screened-Poisson (poisson_solver). */
enum mcpm_long_range {
  MCPM_LR_NONE = 0,
  MCPM_LR_DIRECT = 1,
  MCPM_LR_DENSITY_EXPANSION = 2,
  MCPM_LR_SCREENED_POISSON = 3
};
/* Velocity model of velocity_update (gas.model, Main_Functions.cpp:1545-1600):
   ESMC -> Collision_ESMC_new3, FP_DENSE / FP_IDEAL -> Velocity_FP_parallel
   with or without the dense-gas Y(n) and closures, DSMC_VHS ->
   Collision_DSMC_VHS, NONE -> free flight, DSMC_HS -> Collision_DSMC_HS,
   CBA_VHS -> Collision_CBA_VHS, CBA_HS -> Collision_CBA_HS, FP_LINEAR ->
   Velocity_FP_linear, ESMC_NEW4 -> Collision_ESMC_new4 (slab only; the
   reference's call of it is commented out at Main_Functions.cpp:1560). */
enum mcpm_model {
  MCPM_MODEL_ESMC = 0,
  MCPM_MODEL_FP_DENSE = 1,
  MCPM_MODEL_FP_IDEAL = 2,
  MCPM_MODEL_DSMC_VHS = 3,
  MCPM_MODEL_NONE = 4,
  MCPM_MODEL_DSMC_HS = 5,
  MCPM_MODEL_CBA_VHS = 6,
  MCPM_MODEL_CBA_HS = 7,
  MCPM_MODEL_FP_LINEAR = 8,
  MCPM_MODEL_ESMC_NEW4 = 9
};

/* ---- parameters -------------------------------------------------------- */
/* The subset of the reference's struct GAS / struct BOX that the slab step
   reads (cpp_headers.h:145-269), in plain C types.  Lengths, masses and
   temperatures are in the reference's own units. */
typedef struct mcpm_params {
  int32_t  precision;    /* enum mcpm_precision: particle storage type        */
  int32_t  boundary;     /* enum mcpm_boundary                                 */
  int32_t  y_model;      /* enum mcpm_ymodel                                   */
  int32_t  long_range;   /* enum mcpm_long_range (gas.long_range)              */
  int32_t  n_cells;      /* box.N[1]: cells along the slab axis                */
  int32_t  model;        /* enum mcpm_model (gas.model)                        */
  int64_t  capacity;     /* maximum number of particles the context holds      */
  uint64_t seed;         /* Philox key for every random draw                   */
  double   length;       /* box.Len[1]                                          */
  double   wall_lo;      /* lower diffuse wall (reference: sigma + 1e-15)      */
  double   wall_hi;      /* upper diffuse wall (reference: Len - sigma - 1e-15)*/
  double   T_wall_lo;    /* box.T_wall_1                                        */
  double   T_wall_hi;    /* box.T_wall_2                                        */
  double   U_wall_lo;    /* box.U_wall_1: tangential wall velocity              */
  double   U_wall_hi;    /* box.U_wall_2                                        */
  double   sigma;        /* gas.sigma: hard-sphere diameter                     */
  double   mass;         /* gas.m                                               */
  double   kb;           /* gas.kb                                              */
  double   b;            /* gas.b = 2 pi sigma^3 / 3 (myIO.cpp:2828)            */
  double   Fn;           /* gas.Fn: real molecules per simulated particle       */
  double   cell_volume;  /* cells[c].volume (uniform mesh)                      */
  double   n_ratio;      /* cells[c].n_ratio (1 unless zoomed)                  */
  double   delta_t;      /* gas.delta_t                                         */
  double   ast;          /* gas.ast: Yukawa amplitude (myIO.cpp:2891)           */
  double   bst;          /* gas.bst: Yukawa screening lambda (myIO.cpp:2892)    */
  double   eqdist;       /* gas.eqdist: Dirichlet quadrature reach / 3          */
  double   mu;           /* gas.mu: viscosity at T0 (myIO.cpp:2829)             */
  double   T0;           /* gas.T0: reference temperature of mu(T)              */
  double   visp;         /* gas.visp: mu ~ T^visp (scale_mu, Main_Functions.cpp:1970) */
  double   crref;        /* gas.crref: VHS reference relative speed             */
  double   phi;          /* gas.phi: attraction strength of the direct and
                            density-expansion forces (myIO.cpp:2915-2920)     */
  /* The plane (problem injection, myIO.cpp:2729-2761; FEM.cpp): dims = 2
     adds the x axis.  Then n_cells, length and boundary describe y (box.N[1],
     box.Len[1]) and the fields below describe x.  0 is read as 1. */
  int32_t  dims;         /* 1: slab (the default), 2: plane                     */
  int32_t  n_cells_x;    /* box.N[0]; 2D only                                   */
  int32_t  boundary_x;   /* enum mcpm_boundary along x; 2D only                 */
  int32_t  reserved1;
  double   length_x;     /* box.Len[0]; 2D only                                 */
} mcpm_params;

typedef struct mcpm_ctx mcpm_ctx;

/* Per-step collision statistics (the reference prints the equivalent). */
typedef struct mcpm_collision_stats {
  int64_t candidates;    /* sum over cells of m_c                               */
  int64_t valid;         /* candidates with a partner j != i                    */
  int64_t accepted;      /* collisions performed                                */
  int32_t rounds;        /* parallel rounds the reservation schedule took       */
  int32_t revisits;      /* operations carried into a later round, summed over
                            the rounds (0 where operations run one by one)    */
} mcpm_collision_stats;

/* ---- lifetime ---------------------------------------------------------- */
/* Create a context on the current HIP device.  No reference counterpart:
   replaces the malloc of particle and cell arrays in main.cpp:74-165. */
int mcpm_create(const mcpm_params *p, mcpm_ctx **out);
int mcpm_destroy(mcpm_ctx *ctx);
const char *mcpm_status_string(int status);
/* Block until all work queued on the context's stream has finished. */
int mcpm_sync(mcpm_ctx *ctx);

/* ---- particle data ------------------------------------------------------ */
/* Copy n particles from host SoA arrays (the reference's x2, U1, U2, U3,
   main.cpp:104-165) into the context.  Any order; mcpm_cell_update sorts. */
int mcpm_upload(mcpm_ctx *ctx, int64_t n, const double *y,
                const double *u1, const double *u2, const double *u3);
/* Copy the particles back in the current (cell-sorted) order. */
int mcpm_download(mcpm_ctx *ctx, double *y, double *u1, double *u2, double *u3);
int64_t mcpm_particle_count(const mcpm_ctx *ctx);
/* The plane's particle data: the reference's x1, x2, U1, U2, U3
   (main.cpp:104-165).  mcpm_upload / mcpm_download serve the slab; on a
   plane context mcpm_upload is MCPM_ERR_ARG and mcpm_download leaves x out. */
int mcpm_upload_xy(mcpm_ctx *ctx, int64_t n, const double *x, const double *y,
                   const double *u1, const double *u2, const double *u3);
int mcpm_download_xy(mcpm_ctx *ctx, double *x, double *y,
                     double *u1, double *u2, double *u3);

/* ---- per-step operators -------------------------------------------------- */
/* cell_update (cpp_headers.h:501; Main_Functions.cpp:2138-2749): bin every
   particle into cell floor(y / dy), rebuild the per-cell index lists in
   stable order, and set num_inside and n = n_ratio Fn N_c / V.  Here the
   index lists become a physical stable counting sort of the particles. */
int mcpm_cell_update(mcpm_ctx *ctx);

/* Collision_ESMC_new3 (cpp_headers.h:365; collisions.cpp:583-720): Enskog
   SMC collisions with the Y(n) contact factor.  `step` addresses the random
   stream. */
int mcpm_collide_esmc(mcpm_ctx *ctx, int64_t step, mcpm_collision_stats *stats);

/* Collision_ESMC_new4 (cpp_headers.h:364; collisions.cpp:722-837): the same
   Enskog collisions in new4's own order, cells from the last to the first
   and each cell's candidates in turn, candidates only in cells of two or
   more particles, the partner cell floor((y + sigma dir_y) / dy) clamped into
   the slab (wrapped on a periodic slab), and acceptance only of approaching
   pairs (g.dir > 0).  The strict cell order chains the cells together: the
   reservation rounds grow with the number of cells, up to the schedule's
   limit (MCPM_ERR_CAPACITY beyond it).  Slab only. */
int mcpm_collide_esmc_new4(mcpm_ctx *ctx, int64_t step, mcpm_collision_stats *stats);

/* poisson_solver (cpp_headers.h:450; Main_Functions.cpp:3285-3581): the 1D
   screened-Poisson mean field and the kick U2 += F2 dt.  This call only
   solves for the per-cell force F2; the kick is applied by mcpm_stream_bc
   (or mcpm_kick) in the same pass that moves the particles.
   On a plane context this is compute_SP_2D (cpp_headers.h:398; FEM.cpp:
   249-610): Q1 finite elements on the (N0+1)(N1+1) nodes, stiffness plus
   lambda^2 V mass matrix, right-hand side V M n_n from the nodal density,
   Dirichlet values on the walls from the 50^3-point Yukawa Green's-function
   quadrature (FEM.cpp:196-247, 344-448), core subtraction with its 5-point
   correction (FEM.cpp:497-554); the kick U -= grad(Q1 interpolant) ast / m
   dt (FEM.cpp:568-582) is again left to the push.  The reference factors a
   dense matrix with PARDISO; here conjugate gradients on the lifted system
   solve the same equations. */
int mcpm_poisson_solver(mcpm_ctx *ctx);

/* vlasov_taking_the_integral (cpp_headers.h:628; Main_Functions.cpp:
   3582-3942): face potential V_f = -sum over r in [eqdist, 3 eqdist] and
   polar angle (50 x 50 Gauss-Legendre) of (sigma/r)^6 r^2 2 pi n sin(phi),
   F2 = -(V_{c+1} - V_c) / dy * phi / m.  Like mcpm_poisson_solver it sets F2
   and leaves the kick to the push.
   On a plane context this is compute_Vlasov_2D (cpp_headers.h:390; FEM.cpp:
   613-726): at every node the 100^3-point Gauss-Legendre sum over r in
   [sigma, 3 sigma], azimuth and polar angle of (sigma/r)^6 r^2 n, and the
   kick U -= grad(Q1 interpolant) phi / m dt. */
int mcpm_vlasov_integral(mcpm_ctx *ctx);

/* The plane's node arrays, (N0+1)(N1+1) doubles each, node (i, j) at
   j (N0+1) + i (cells[c].id, myIO.cpp:1447-1451); either pointer may be NULL.
   n_node: the nodal density deposit of cell_update (Main_Functions.cpp:
   2268-2364): Fn / V per particle in the node's dual cell, doubled on each
   wall the node lies on; computed by this call.  potential: the nodal
   field x of the last mcpm_poisson_solver / mcpm_vlasov_integral. */
int mcpm_get_nodes(mcpm_ctx *ctx, double *n_node, double *potential);

/* Density-expansion attraction (long_range == 2, Main_Functions.cpp:
   1632-1653) from the Gaussian kernel-density gradient of cell_update
   (Main_Functions.cpp:2378-2427): F2 = -dn2 4 pi (2^(1/6) sigma)^3 2 phi / m
   in every cell but the two wall cells.  Sets F2. */
int mcpm_density_expansion(mcpm_ctx *ctx);

/* The kernel density estimate of cell_update (Main_Functions.cpp:2378-2427):
   n_smooth, dn2, dn3 per cell (h = dy/2, cut-off 2 dy).  Any pointer may be
   NULL; computes them if no call has since the last sort. */
int mcpm_get_kde(mcpm_ctx *ctx, double *n_smooth, double *dn2, double *dn3);

/* The velocity kick of poisson_solver (Main_Functions.cpp:3574-3577) alone,
   U2 += F2[cell] dt, for callers that want the reference's operator split. */
int mcpm_kick(mcpm_ctx *ctx);

/* stream_particles + apply_BC (cpp_headers.h:495,498; Main_Functions.cpp:
   1451-1521, 4917-5248; Fokker_Planck.cpp:948-969), followed by the binning
   of cell_update: y += U2 dt, wall re-emission or periodic wrap, re-sort.
   When `with_kick` is non-zero the pending F2 kick is applied first, fused
   into the same pass. */
int mcpm_stream_bc(mcpm_ctx *ctx, int64_t step, int32_t with_kick);

/* One whole time step in the reference's order (main.cpp:389-669): the
   velocity model of params.model (ESMC collisions, FP closures + velocity
   update, DSMC collisions or nothing), mean-field solve, kick, the FP_dense
   position drift (DFP_extra_streaming, Fokker_Planck.cpp:904-946), stream,
   boundary and re-sort. */
int mcpm_step(mcpm_ctx *ctx, int64_t step, mcpm_collision_stats *stats);

/* ---- host-free steps (no reference counterpart: main.cpp:389-669 calls the
   step's functions one by one from the host) ------------------------------
   mcpm_graph_capture records one mcpm_step of the context as HIP graphs
   (one per parity of the particle buffers): every launch reads the step
   number, the candidate count and the round tags from device memory, the
   reservation rounds are `rounds` grid rounds (0: the band count + 6)
   followed by one workgroup that finishes what is left, and nothing waits
   for the host inside a step.  Covers ESMC (Collision_ESMC_new3) on a slab
   or a batch of slabs, without a mean field or with the screened-Poisson
   one, sorted by the windowed sort, without pressure sampling, particle
   ids, per-cell weights or the debug log (MCPM_ERR_ARG for other models,
   MCPM_ERR_STATE when the particles are not binned or the windowed sort
   does not apply).  `op_capacity` bounds the candidates of one step (0:
   twice the last collision call's, at least N).
   mcpm_graph_steps runs steps step0 .. step0 + nsteps - 1 by replaying the
   graphs; the results are bit for bit those of as many mcpm_step calls.
   It looks at the device state after 8 steps, then after twice as many
   each time up to 64, and after 8 again following a stop (the replays
   queued behind a stopped step run as no-ops).  A step whose candidates
   exceed the capacity is recaptured with twice the capacity and run again;
   a step whose particle left its sort window finishes its sort on the host
   (full histogram), the windows widen (the graphs read their width from
   device memory) and the same graphs replay on; stats are sums over
   the steps (rounds: summed; revisits: operations the one-workgroup tail
   handled).  Allocates only when it captures. */
int mcpm_graph_capture(mcpm_ctx *ctx, int64_t op_capacity, int32_t rounds);
int mcpm_graph_steps(mcpm_ctx *ctx, int64_t step0, int64_t nsteps,
                     mcpm_collision_stats *stats);
int mcpm_graph_release(mcpm_ctx *ctx);

/* ---- cell state ---------------------------------------------------------- */
/* omega_max initialisation of main.cpp:250-269: crm = 10 sqrt(kb T_c / m),
   omega_max = 4 pi sigma^2 crm n_c dt, times Y(n_gas) for ESMC.  T_c is
   taken from per-cell moments; `n_gas` is gas.n. */
int mcpm_init_omega_max(mcpm_ctx *ctx, double n_gas, int32_t times_y);
int mcpm_set_omega_max(mcpm_ctx *ctx, const double *omega_max);

/* Cell arrays in cell order; any pointer may be NULL.  count = num_inside,
   start = offset of the cell's first particle, n = number density,
   omega_max, F2 = mean-field force per unit mass on the slab axis. */
int mcpm_get_cells(mcpm_ctx *ctx, int32_t *count, int32_t *start, double *n,
                   double *omega_max, double *F2);

/* Per-cell moments of cell_update (Main_Functions.cpp:2461-2593), FP64,
   two-pass: out[c*MCPM_NMOM + k] with k = 0..2 U_space, 3..8 PIJ (11,12,
   13,22,23,33), 9..18 M3 (111,112,113,122,123,133,222,223,233,333),
   19..24 M4, 25..27 M5, 28 DM2, 29 T, 30 N_c, 31..33 Mp (mean peculiar
   velocity, zero up to round-off).  out = NULL takes the moments on the
   device only (as cell_update does every step), without a copy or a wait. */
#define MCPM_NMOM 34
int mcpm_moments(mcpm_ctx *ctx, double *out);

/* ---- Fokker-Planck model ------------------------------------------------- */
/* The closure part of cell_update for the FP models: Lambda and gamma_st
   (Main_Functions.cpp:2602-2605, 2718-2728), alpha_calculation when model is
   FP_DENSE and step > 0 (cpp_headers.h:541; Fokker_Planck.cpp:200-587),
   c_gamma_calculation (cpp_headers.h:467; Fokker_Planck.cpp:590-804), both
   with Gauss_Jordan (Math_Func.cpp:165-232), and phi = alpha Mp
   (Main_Functions.cpp:2734-2746).  Needs sorted particles. */
int mcpm_fp_closures(mcpm_ctx *ctx, int64_t step);
/* random_generator + Velocity_FP_parallel (cpp_headers.h:362,632;
   Fokker_Planck.cpp:20-145, 1065-1182): per-cell zero-mean unit-variance
   Gaussian noise, exact Ornstein-Uhlenbeck step with the closures' drift,
   per-cell kinetic-energy rescale.  Uses the last mcpm_fp_closures.  With
   model FP_LINEAR this is Velocity_FP_linear (cpp_headers.h:517;
   Fokker_Planck.cpp:1186-1287): the same noise and rescale, relaxation
   time tau = 2 mu(T) / (n kB T) and no drift (of the closures only tau is
   read). */
int mcpm_velocity_fp(mcpm_ctx *ctx, int64_t step);
/* Closures in cell order: out[c*MCPM_NCLOS + k], k = 0 Lambda, 1..6 c_ij,
   7..9 gamma_i, 10..18 alpha, 19..21 beta, 22 gamma_st, 23..25 phi,
   26 div U, 27 tau = 2 mu(T) / (n kB T). */
#define MCPM_NCLOS 28
int mcpm_get_closures(mcpm_ctx *ctx, double *out);

/* ---- plain DSMC ------------------------------------------------------------ */
/* Collision_DSMC_VHS (cpp_headers.h:612; collisions.cpp:2063-2141): NTC
   pairs inside each cell, VHS cross-section, isotropic scattering, per-cell
   crm (kept in the omega_max array) <- max(1.1 crm_seen, 0.9 crm). */
int mcpm_collide_dsmc_vhs(mcpm_ctx *ctx, int64_t step, mcpm_collision_stats *stats);
/* Collision_DSMC_HS (cpp_headers.h:507; collisions.cpp:1989-2061): as
   DSMC_VHS with the hard-sphere weight sigma_T = Vr, and crm updated in every
   cell.  Cells with fewer than two particles draw no candidates. */
int mcpm_collide_dsmc_hs(mcpm_ctx *ctx, int64_t step, mcpm_collision_stats *stats);
/* Collision_CBA_VHS (cpp_headers.h:611; collisions.cpp:2144-2248): DSMC_VHS
   with the candidate count scaled by the contact value Y(n) of the cell, and
   each accepted pair displaced by +-sigma along the change of its relative
   velocity (x1 and x2 of the reference; the slab coordinate is x2). */
int mcpm_collide_cba_vhs(mcpm_ctx *ctx, int64_t step, mcpm_collision_stats *stats);
/* Collision_CBA_HS (cpp_headers.h:610; collisions.cpp:2251-2346): per cell,
   Vr_max = sqrt(pi kB T / m), M = Fn N^2 pi sigma^2 Vr_max dt / (2 V_c) times
   (1 - b rho / 8) / (1 - b rho / 4)^3 (rho = n m) candidates, drawn with
   replacement and run in order; an accepted pair faster than Vr_max raises
   Vr_max and with it the number of candidates still to run.  Displacement
   as CBA_VHS.  Cells run in parallel, the candidates of a cell one after
   the other as in the reference.  MCPM_ERR_CAPACITY when a cell's bound
   passes 2^26 candidates (non-finite or absurd velocities). */
int mcpm_collide_cba_hs(mcpm_ctx *ctx, int64_t step, mcpm_collision_stats *stats);

/* cell_poly_velocity (cpp_headers.h:470; Main_Functions.cpp:1918-1968):
   accumulate the per-cell time-average sums.  mcpm_get_samples returns
   out[c*MCPM_NSAMP + k]: 0 sum_weight, 1..3 sum_M, 4..9 sum_MM, 10..19
   sum_MMM (m,n,l order of the reference), 20..25 sum_MMMM. */
#define MCPM_NSAMP 26
int mcpm_sample(mcpm_ctx *ctx);
int mcpm_get_samples(mcpm_ctx *ctx, double *out);
int mcpm_reset_samples(mcpm_ctx *ctx);

/* measure_pressure (cpp_headers.h:388; Main_Functions.cpp:13-34, called
   after streaming at main.cpp:490-492): while enabled, every moving push
   (mcpm_stream_bc, mcpm_step) adds Fn m |U2| to p_m of each cell whose upper
   face lies between the particle's cell before streaming (after the FP drift)
   and its cell after the boundary.  On a periodic slab the faces of the
   unwrapped flight are taken.  Slab only.  p_m accumulates on the device
   until mcpm_reset_samples. */
int mcpm_measure_pressure(mcpm_ctx *ctx, int32_t enable);

/* Particles re-inserted by the REINJECT boundary since the context was
   created (the reference's box.out of BC_evaporation, which it counts once
   step > box.after; here every step). */
int mcpm_get_escaped(mcpm_ctx *ctx, int64_t *count);

/* Energy check, FP64 on the device whatever the particle precision: out[0]
   = (1/2) m Fn sum |U|^2 (kinetic energy of the molecules the particles
   stand for), out[1..3] = m Fn sum U (momentum).  Not on batched replicas. */
int mcpm_energy_momentum(mcpm_ctx *ctx, double *out);

/* The energy check without allocation.  No counterpart in the reference:
   for callers that cannot allocate device memory (code under stream
   capture, tight serving loops, harnesses that forbid hipMalloc).  The
   same sums, in the same order, as mcpm_energy_momentum, over caller-owned
   device arrays u1, u2, u3 of n FP64 velocities; out: device pointer to 4
   doubles, out[0] = (1/2) mass_fn sum |U|^2, out[1..3] = mass_fn sum U
   (mass_fn = m Fn).  Asynchronous on `stream` (a hipStream_t; null for the
   default stream).  Calls no hipMalloc/hipFree: the per-workgroup partials
   sit in a fixed 32 KiB device array of the library (one per GPU, kept
   until the process ends), which holds them for any n.  Calls that overlap
   on one GPU share that array and race.  MCPM_ERR_ARG for n < 0 or a null
   pointer. */
int mcpm_energy_momentum_noalloc(const double *u1, const double *u2, const double *u3,
                                 int64_t n, double mass_fn, double *out, void *stream);

/* Particle identities.  The reference keeps every particle at its own index
   of the caller's arrays and rebuilds cells[c].indices_inside from those
   indices (Main_Functions.cpp:2367-2376); here particles move through the
   stable sort.  With tracking on, every sort also carries each particle's
   upload slot (its index in the arrays of the last mcpm_upload /
   mcpm_upload_xy / mcpm_initialize, or in the order when tracking was
   switched on), and mcpm_download_ids returns, for each position of
   mcpm_download, that slot: a caller can keep its own indexing. */
int mcpm_track_ids(mcpm_ctx *ctx, int32_t enable);
int mcpm_download_ids(mcpm_ctx *ctx, int32_t *ids);

/* Particle weights that differ from cell to cell (the inverted and
   evaporation problems' heavy liquid zones, myIO.cpp:1493-1525).
   mcpm_set_cell_ratio: cells[num].n_ratio for every cell (n_cells integers
   >= 1), which then enters each cell's density n = n_ratio Fn N / V
   (Main_Functions.cpp:2588); null restores p.n_ratio everywhere.  Every
   particle takes its cell's weight (main.cpp:239-242), now if the particles
   are sorted, else at the next sort (and again after an upload or
   initialisation); the sorts carry the weights with the particles.
   mcpm_set_particle_ratio / mcpm_get_particle_ratio: the weights of the
   particles in mcpm_download order.  Slab contexts without replicas. */
int mcpm_set_cell_ratio(mcpm_ctx *ctx, const int32_t *ratio);
int mcpm_set_particle_ratio(mcpm_ctx *ctx, const int32_t *w);
int mcpm_get_particle_ratio(mcpm_ctx *ctx, int32_t *w);

/* inverted_addremove_particles (Main_Functions.cpp:1206-1450) with gas
   n_ratio = n_ratio on the sorted particles (after mcpm_cell_update or a
   step): a particle heavier than its cell splits into n_ratio particles (its
   n_ratio - 1 copies at its position, velocity + h N(0,1) per component,
   h = 1.06 sig (n_rem + n_ratio)^(-1/5), sig and the mean velocity formed as
   at :1302-1325), a lighter one is deleted with probability 1 - 1/n_ratio or
   takes the cell's weight; then every particle has its cell's weight.  N
   changes; the particles stay sorted, each cell's kept particles first and
   its new ones after them.  Philox draws (p, copy, step).  MCPM_ERR_CAPACITY
   (nothing changed) when the particles would exceed the capacity;
   MCPM_ERR_STATE without cell weights or unsorted particles; MCPM_ERR_ARG
   with identity tracking on. */
typedef struct mcpm_addremove_stats {
  int64_t before;   /* particles before */
  int64_t split;    /* particles that split */
  int64_t created;  /* new particles: split x (n_ratio - 1) */
  int64_t deleted;  /* particles removed */
  int64_t after;    /* particles after */
} mcpm_addremove_stats;
int mcpm_addremove(mcpm_ctx *ctx, int32_t n_ratio, int64_t step, mcpm_addremove_stats *stats);

/* thermostat (Main_Functions.cpp:1674-1869), the inverted problem's branch
   (:1760-1830): after the cell moments are taken from the current particles
   (cell_update), every particle of a cell whose centre lies below zone_lo
   (box.thermc) has its peculiar velocity scaled by sqrt(T_lo / T_cell)
   (gas.Tc), and every particle of a cell whose centre lies above
   length - zone_hi (box.thermh) by sqrt(T_hi / T_cell) (gas.Th); cells with
   T_cell = 0 are left alone.  Each cell keeps its mean velocity.  Slab
   only. */
int mcpm_thermostat(mcpm_ctx *ctx, double zone_lo, double T_lo, double zone_hi, double T_hi);
/* The pressure columns of PostProc (myIO.cpp:743-744, 855, 889-890): p_m,
   p_rep = n kB T (1 + n b Y(n)) and ptot = p_rep - 2 b n^2 phi, the last two
   from the current particles (cell T of the moments, cell n).  Arrays of
   n_cells doubles; any of them may be null.  Slab only. */
int mcpm_get_pressure(mcpm_ctx *ctx, double *p_m, double *p_rep, double *p_tot);

/* increase_collision_rate (cpp_headers.h:505; Main_Functions.cpp:1906-1915),
   evaluated on the host with the context's parameters. */
double mcpm_contact_y(const mcpm_params *p, double n);

/* ---- batched replicas ------------------------------------------------------ */
/* No reference counterpart (the reference runs one system per process): many
   independent slabs of one length in one context, laid end to end in its
   cells and particles, run by the same launches.  Each replica has its own
   cells, cell volume, Fn and seed and draws exactly what it would draw as a
   context of its own, so its particles evolve bit for bit as in an unbatched
   run.  Models ESMC and NONE; no field or the screened-Poisson field; all
   three slab boundaries.  The shared constants come from `common`, whose
   n_cells and capacity are ignored. */
typedef struct mcpm_replica {
  int32_t  n_cells;      /* cells of this replica                               */
  int32_t  reserved;
  int64_t  capacity;     /* most particles it holds                             */
  uint64_t seed;         /* its Philox key                                      */
  double   cell_volume;  /* its cell volume                                     */
  double   Fn;           /* its particle weight                                 */
} mcpm_replica;
int mcpm_create_batch(const mcpm_params *common, int32_t n_rep, const mcpm_replica *reps,
                      mcpm_ctx **out);
/* Particles replica after replica (counts[r] of them each); the counts stay
   fixed afterwards.  mcpm_download returns them in the same replica order,
   each replica's in its cell order.  mcpm_upload is refused on a batch. */
int mcpm_upload_batch(mcpm_ctx *ctx, const int64_t *counts, const double *y,
                      const double *u1, const double *u2, const double *u3);
/* mcpm_init_omega_max with each replica's own mean density n_gas[r]. */
int mcpm_init_omega_max_batch(mcpm_ctx *ctx, const double *n_gas, int32_t times_y);

/* ---- settings file, initial state and PostProc output ---------------------- */
/* What readSettingsFile (cpp_headers.h:624; myIO.cpp:1756-3016) derives from a
   reference settings file, for the problems on this path: closed_box (1,
   slab, specular walls), couette_flow (2, slab, diffuse walls at sigma and
   L - sigma) and injection (12, plane, specular walls).  params is ready for
   mcpm_create (seed 1, FP64, the reference's Y(n); set them afterwards). */
typedef struct mcpm_settings {
  mcpm_params params;
  int64_t  n_particles;      /* gas.N after the problem's set-up (injection adds Nh) */
  int64_t  n_drop;           /* injection: gas.Nh, particles of the liquid disc */
  int32_t  problem;          /* box.problem */
  int32_t  num_steps;        /* box.num_steps ("steps") */
  int32_t  every;            /* box.every: PostProc cadence */
  int32_t  after;            /* box.after: sampling starts after step times * after */
  int32_t  reset_every;      /* box.reset_every */
  int32_t  num_realizations; /* box.num_realizations */
  int32_t  init_step;        /* box.init_step: first step with collisions */
  int32_t  direction[3];     /* box.direction: axes that stream */
  int32_t  num_cells[3];     /* box.N */
  int32_t  reserved2;
  double   box_len[3];       /* box.Len after the problem's scaling */
  double   n;                /* gas.n */
  double   T;                /* gas.T */
  double   U0[3];            /* gas.U0: mean of the initial Maxwellian */
  double   T_init;           /* temperature of the initial Maxwellian (Main_Functions.cpp:4791-4792) */
  double   drop_dia;         /* injection: box.drop_dia (scaled by sigma) */
  double   epsilon;          /* gas.epsilon */
  double   lambda;           /* gas.lambda: mean free path (myIO.cpp:2837-2840) */
  double   Kn;               /* gas.Kn */
  char     model[32];        /* gas.model as written in the file */
  char     output[128];      /* gas.output_name */
} mcpm_settings;

/* Parse a settings file: "keyword values..." lines, '#' in the first column
   starts a comment, the rest of a line after its values is ignored; an
   unknown keyword is MCPM_ERR_ARG (the reference aborts).  MCPM_ERR_ARG also
   for problems and models this library does not run. */
int mcpm_read_settings(const char *path, mcpm_settings *out);

/* initialization (cpp_headers.h:566; Main_Functions.cpp:4222-4815) for the
   problems above, drawn on the device with Philox (stream MCPM_TAG_INIT,
   particle index i): positions uniform along the streaming axes (the others
   at L/2); injection: the first N - Nh uniform in the plane, the last Nh in
   the disc of diameter drop_dia at the centre (r = R sqrt(u), theta = 2 pi
   u'); velocities Maxwellian about U0 at T_init.  Replaces the particles of
   the context (n_particles of them), then bins them. */
int mcpm_initialize(mcpm_ctx *ctx, const mcpm_settings *s);

/* write_post_processing's PostProc file (cpp_headers.h:399; myIO.cpp:597-897): dir/<output>_
   <direct|denexp|screened-Poisson|noatt>_<step>.dat, the reference's header
   line and 118 columns per cell, one block per x column.  Columns this path
   computes: cell geometry, t_avg (dt times the samples taken since the last
   reset), dt, U_space, T, n, the cell_poly_velocity sums, volume, n_smooth,
   N, ptot, Fn, F2, PIJ[3], Q[1], div U, alpha, beta, gamma_st, phi, p_m,
   p_rep; the others (face sums of the inverted problem, xx, UU*) are 0. */
int mcpm_write_postproc(mcpm_ctx *ctx, const mcpm_settings *s, int64_t step, const char *dir);

/* ---- diagnostics --------------------------------------------------------- */
/* Record the outcome of every valid collision candidate of the next
   mcpm_collide_esmc call (for parity checks): priority, i, j, accepted. */
int mcpm_debug_collision_log(mcpm_ctx *ctx, int32_t enable);
int mcpm_get_collision_log(mcpm_ctx *ctx, int64_t cap, uint32_t *prio, uint32_t *i,
                           uint32_t *j, uint8_t *accepted, int64_t *n_out);

/* Average device duration (ms) of each kernel class over the calls since the
   last reset, measured with HIP events on the context's stream.  `which`:
   0 collide rounds, 1 sort (push+scatter), 2 candidate generation, 3 field,
   4 whole step, 5 Fokker-Planck velocity update, 6 Fokker-Planck moments and
   closures. */
int mcpm_timing_enable(mcpm_ctx *ctx, int32_t enable);
int mcpm_timing_get(mcpm_ctx *ctx, int32_t which, double *avg_ms, int64_t *calls);

/* ---- device random-access rates ------------------------------------------- */
/* No counterpart in the reference: a measurement of the current device for
   the roofline of the reservation rounds, whose accesses all go to
   unpredictable particles.  Over n_records records of record_bytes (16 or 32),
   n_records 4-byte words and n_records / 32 bytes (the rounds' marks, one
   byte per 32 particles), with indices hashed uniformly,
   the best of five launches of min(n_records, 2^24) accesses each, in
   accesses per second: record gathers, word gathers, 32-bit atomic minimums
   (result unused), record stores, word stores and byte gathers.  Allocates
   and frees its own arrays. */
typedef struct mcpm_access_rates {
  double gather_record;
  double gather_word;
  double atomic_min_word;
  double scatter_record;
  double scatter_word;
  double gather_byte;
} mcpm_access_rates;
int mcpm_probe_random_access(int64_t n_records, int32_t record_bytes, mcpm_access_rates *out);

/* ---- launch-shape knobs ---------------------------------------------------- */
/* No counterpart in the reference (its loops are serial): the tile and
   launch sizes of the GPU schedule, for an autotuner.  None of them changes
   any result, only the speed.
   MCPM_TUNE_BANDS       priority bands of the ESMC reservation rounds, 1..16
   MCPM_TUNE_ROUND_GRID  most workgroups of one reservation round (multiple of 8)
   MCPM_TUNE_SORT_CHUNK  particles per sort chunk (multiple of 256, 1024..32768;
                         0 = about 2048 chunks)
   MCPM_TUNE_SORT_WINDOW 1 (default): on slabs whose particles moved a few
                         cells, sort by per-chunk windowed histograms (1024
                         particles per chunk, a few cells each), redone by the
                         full histogram when a particle left its window; 0: the
                         full histogram always
   MCPM_TUNE_SORT_WINDOW_RUNS, MCPM_TUNE_SORT_WINDOW_REDONE (read only):
                         windowed sorts run and redone since creation
   MCPM_TUNE_ROUND_UNROLL operations per thread of a reservation round, 2 or
                         4; 0 (default): 4 when a priority band's operations
                         fit in one pass of the round grid at 2 per thread,
                         else 2 (DSMC: 4)
   A captured step (mcpm_graph_capture) keeps the values it was captured
   with; capture again for new ones. */
#define MCPM_TUNE_BANDS      0
#define MCPM_TUNE_ROUND_GRID 1
#define MCPM_TUNE_SORT_CHUNK 2
#define MCPM_TUNE_SORT_WINDOW 3
#define MCPM_TUNE_SORT_WINDOW_RUNS 4
#define MCPM_TUNE_SORT_WINDOW_REDONE 5
#define MCPM_TUNE_ROUND_UNROLL 6
int mcpm_set_tuning(mcpm_ctx *ctx, int32_t knob, int64_t value);
int mcpm_get_tuning(mcpm_ctx *ctx, int32_t knob, int64_t *value);

/* ---- learned field surrogate (spectral neural operator) ------------------ */
/* No counterpart in the reference: a network that stands in for the periodic
   plane's screened-Poisson solve (compute_SP_2D, FEM.cpp:249-610).  It maps
   the nodal density of an N0 x N1 periodic plane (unique nodes, p = j N0 + i)
   to the Q1 FEM solution of (K + lambda^2 M) phi = M n before the core
   subtraction (FEM.cpp:497-554).  Width 64, four Fourier layers with `modes`
   retained modes per axis (8..32, multiple of 8; ky in [0,m) and [N1-m,N1),
   kx in [0,m)), tanh GELU, BF16 activations on MFMA with FP32 accumulation.
   The flat FP32 parameter vector has mcpm_fno_param_count entries:
   lift w (64), b (64); per layer W[o][i] (64 x 64), b (64), spectral
   weights R[a][kx][i][o] complex (re, im) (2m x m x 64 x 64 x 2); projection
   w (64), b (1); zero padding to a multiple of 4.
   Normalisation: u = (n - in_shift) in_scale, phi = y / out_scale + out_shift.
   Loss over a snapshot (e = y - t in normalised units, periodic forward
   differences): sum(|grad e|^2 + w0 e^2) / sum(|grad t|^2 + w0 t^2).
   N0 must be a multiple of 64, N1 a multiple of 16 and at least 2 modes.
   The object owns its stream and device memory; calls are synchronous
   except where they run on a context's stream (add_snapshot, field). */
typedef struct mcpm_fno mcpm_fno;
typedef struct mcpm_fno_field_stats {
  double residual;        /* |M n - A phi_net| / |M n| of the network's field */
  int32_t cg_iterations;  /* conjugate-gradient iterations that followed (0 when the
                             exact solve was the FFT of a power-of-two plane) */
  int32_t used_network;   /* 1 when the network's field was kept as it is */
} mcpm_fno_field_stats;
int mcpm_fno_param_count(int32_t modes, int64_t *count);
int mcpm_fno_create(int32_t n0, int32_t n1, int32_t modes, const float *params, mcpm_fno **out);
int mcpm_fno_destroy(mcpm_fno *fno);
int mcpm_fno_set_params(mcpm_fno *fno, const float *params);
int mcpm_fno_get_params(mcpm_fno *fno, float *params);
int mcpm_fno_set_norm(mcpm_fno *fno, double in_shift, double in_scale, double out_shift,
                      double out_scale);
/* Inference on host arrays of N0 N1 normalised densities / outputs. */
int mcpm_fno_forward(mcpm_fno *fno, const float *u, float *y);
/* Loss and full parameter gradient of one sample (host arrays; grad and y
   may be null). */
int mcpm_fno_loss_grad(mcpm_fno *fno, const float *u, const float *t, float w0, float *loss,
                       float *grad, float *y);
/* Snapshot store on the device (normalised density and target).  Every
   snapshot uses the same w0.  mcpm_fno_add_snapshot solves the context's
   field exactly (conjugate gradients from zero) and stores density and
   solution; the context's field is then the exact one, as after
   mcpm_poisson_solver. */
int mcpm_fno_reserve_snapshots(mcpm_fno *fno, int32_t capacity);
int mcpm_fno_add_snapshot_host(mcpm_fno *fno, const float *u, const float *t, float w0);
int mcpm_fno_add_snapshot(mcpm_fno *fno, mcpm_ctx *ctx, float w0);
int mcpm_fno_snapshot_count(mcpm_fno *fno, int32_t *count);
/* `iters` Adam steps (beta 0.9, 0.999, eps 1e-8) on the device, one snapshot
   per step drawn from the first n_snap (all when <= 0) by a splitmix64
   stream of `seed`; losses (may be null) gets each step's loss. */
int mcpm_fno_train(mcpm_fno *fno, int32_t iters, float lr, uint64_t seed, int32_t n_snap,
                   float *losses);
/* The context's field from the network: nodal density, network, the field
   written as the solution; then, with exact = 1, conjugate gradients from it
   to the context's tolerance; with accept_residual >= 0, the same unless
   the network's residual is at most accept_residual; with accept_residual
   < 0, the network's field as it is.  Then the core subtraction and the
   kick of the next push, as mcpm_poisson_solver. */
int mcpm_fno_field(mcpm_fno *fno, mcpm_ctx *ctx, double accept_residual, int32_t exact,
                   mcpm_fno_field_stats *stats);
/* Attach (fno) or detach (null) the surrogate for mcpm_step and
   mcpm_poisson_solver on a periodic plane: the step's field is exact (from
   the network's field) when step % every == 0 (every > 0), else the
   network's field when its residual is at most accept_residual (always
   when accept_residual < 0), else exact.  Detach before destroying fno. */
int mcpm_set_field_surrogate(mcpm_ctx *ctx, mcpm_fno *fno, double accept_residual, int64_t every);
int mcpm_get_field_surrogate_stats(mcpm_ctx *ctx, mcpm_fno_field_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* MCPM_H */
