/*
 * mcpm_dropin.h — the reference program's own C++ call surface, served by
 * libmcpm_dropin.so on top of libmcpm.so.
 *
 * mohsensadr/Monte-Carlo-Particle-Method declares its per-step operators as
 * C++ functions over three structs, GAS, BOX and CELLS (src/cpp_headers.h:
 * 145-347 and 360-730), and its time loop (main.cpp:389-669) calls them on
 * caller-owned arrays of doubles.  This header declares the same structs
 * with the same members in the same order and of the same types, so that
 * objects built by the reference's own code have the layout assumed here,
 * and the same functions with the same signatures, so that their mangled
 * names are those the reference's objects reference.  A reference build
 * links libmcpm_dropin.so in place of collisions.cpp, Fokker_Planck.cpp and
 * FEM.cpp (and of the Main_Functions.cpp definitions listed below, which
 * the executable's own definitions would otherwise take precedence over).
 *
 * The sizes and member offsets asserted at the end are those of the
 * reference's structs under g++ 11 on x86-64 (libstdc++ std::string).
 *
 * Behaviour: see the comment above each function.  All functions keep the
 * caller's particle order: the library tracks each particle's index, writes
 * the particles back to the caller's arrays at their own indices, and
 * rebuilds index[] and cells[c].indices_inside / num_inside after every
 * operation that moves particles between cells (the lists hold the same
 * particles as the reference's, in the library's stable order).  Errors
 * that the reference answers with printf + exit (or does not check) throw
 * mcpm_dropin_error, whose code is an MCPM_ERR_* status.
 */
#ifndef MCPM_DROPIN_H
#define MCPM_DROPIN_H

#include <stdexcept>
#include <string>

#include "mcpm.h"

#define MCPM_REF_NSD 3 /* the reference's NSD (cpp_headers.h:24) */

/* ---- the reference's problem numbers used here (cpp_headers.h:26-37) ---- */
enum mcpm_ref_problem {
  MCPM_REF_CLOSED_BOX = 1,
  MCPM_REF_COUETTE_FLOW = 2,
  MCPM_REF_INJECTION = 12
};

/* struct GAS (cpp_headers.h:145-204): gas and run parameters. */
struct GAS {
  int gp_nn;
  double r_cut, h, mb;
  double N_abs;
  double Lc_ratio, Lh_ratio, Lc_center, Lh_center;
  double sigc, alphac, sigh, alphah;
  double count_SP_c, count_SP_h, count_reem_c, count_reem_h, count_evap_c, count_evap_h;
  double nvc, nvh;
  double nc, nh;
  double nv1, nv2, nv3;
  double Tc, Th, Tv1, Tv2, Tv3;
  double Lc, Lh, Lv, Lv1, Lv3;
  int LcN, LhN, LvN, LvN1, LvN3;
  int Nc1, Nc2, Nc3, Nh1, Nh2, Nh3;
  int Nc, Nh;
  int Nv1, Nv2, Nv3;
  int N13;
  double phi;     /* attraction strength of the direct integral */
  double eqdist;  /* Dirichlet quadrature reach / 3 */
  int n_inflow;
  double ast, bst; /* Yukawa amplitude and screening */
  double sigma;
  double n;
  long int N;      /* number of particles */
  double w;
  double m;
  double lambda;
  double Kn;
  double T;
  double T0;
  double delta_t;
  double frac_mft;
  double kb;
  double U0[MCPM_REF_NSD];
  double U[MCPM_REF_NSD];
  double u_avg;
  double mu;
  double mu_corr;
  double M[MCPM_REF_NSD * 2];
  std::string model; /* "ESMC", "FP_dense", "FP_ideal_gas", "FP_linear", ... */
  std::string output_name;
  double p;
  double Fn, Fn_SPH;
  double factor;
  double phi0;
  double b;
  double tau_shear[9];
  double nualpha;
  double visp;
  double crref;
  double t_post;
  double avgeraging_time_till_now;
  double times;
  double epsilon;
  int long_range;  /* 0 none, 1 direct integral, 2 density expansion, 3 screened Poisson */
  int n_ratio;
  double k;
};

/* struct BOX (cpp_headers.h:206-269): geometry and run control. */
struct BOX {
  double drop_dia;
  double x2_prob[4];
  double ghost, thermc, thermh;
  double JR[4], JL[4], Jref[4];
  double NR[4], NL[4];
  double L1[2], L2[2];
  int N1[2], N2[2];
  double s1, s2;
  double dcones_n;
  double dy0, dx0;
  double dy1, dx1;
  double alpha[2];
  double Lx[3];
  double Ly;
  double s;
  int Ny;
  int Nx[3];
  int N_grid[3];
  double Len[MCPM_REF_NSD];        /* box lengths */
  double delta_dim[MCPM_REF_NSD];  /* cell sizes */
  double Area[MCPM_REF_NSD * 2];
  double area, dummy_nom;
  double Volume;
  int N[MCPM_REF_NSD];             /* cells per direction */
  int num_steps;
  int num_realizations;
  int movie;
  int every;
  int after;
  int reset_every;
  int init_step;
  int num_thr;
  int problem;
  double U_wall_1, U_wall_2, U_wall_3, U_wall_4, U_wall_5, U_wall_6, U_wall_7;
  double T_wall_1, T_wall_2, T_wall_3, T_wall_4, T_wall_5, T_wall_6, T_wall_7;
  double n_wall_4, n_wall_5, n_wall_6, n_wall_7;
  int Ndot_w7, Ndot_w6, Ndot_w5, Ndot_w4;
  int N_tot_genenrate;
  std::string file_name;
  int reset;
  int direction[3];  /* axes that stream */
  int step;          /* the current step: addresses the random streams */
  double out;
};

/* struct CELLS (cpp_headers.h:271-347): per-cell state, one per cell. */
struct CELLS {
  double p_m;
  double UU, UU0, Force;
  double UUR, UU0R;
  double divU;
  int model;
  int *id_split_part;
  int n_ratio;
  int id[4];
  double n_n[4];   /* nodal density copies (plane) */
  double x_n[4];
  double y_n[4];
  int in;
  int ng[2];
  double V[4];
  double dx, dy;
  double dn2;
  double dn3;
  double n_smooth;
  double weight;   /* particles of the cell */
  double cell_center[3];
  double dim[MCPM_REF_NSD * 2];
  double volume;
  int *indices_inside;  /* particles of the cell (malloc'd) */
  int num_inside;
  double alpha[9];
  double sum_alpha[9];
  double beta[3];
  double U_space[3];  /* mean velocity */
  double T;
  double n;
  double U[3];
  double PIJ[6];
  double M4[6];
  double M5[3];
  double Q[3];
  double M3[10];
  double DM2;
  double DM4;
  double c[6];
  double gamma[3];
  double gamma_st;
  double crm;          /* DSMC: the largest relative speed */
  double M_n[6];
  double sum_M_n[6];
  double sum_M_np[6];
  double T_f[6];
  double M1_f[6];
  double M2_f[6];
  double M3_f[6];
  double sum_T_f[4];
  double sum_M1_f[6];
  double sum_N[4];
  double sum_M2_f[6];
  double sum_M3_f[6];
  double sum_MM_f_0[6];
  double sum_MM_f_1[6];
  double sum_MM_f_2[6];
  double sum_MM_f_3[6];
  double sum_MMM[10];
  double sum_MMMM[6];
  double sum_MM[6];
  double sum_M[3];
  double sum_weight;
  double Mp[3];
  double phi[3];
  double omega_max;
  double Lambda;
  double G_sum;
  double F1, F2;       /* mean-field force of the cell */
  double xx;
};

/* The reference's Collision scratch (cpp_headers.h:350-357); only passed
   through here. */
struct Collision;

/* An MCPM_ERR_* status from the library (see mcpm_status_string). */
class mcpm_dropin_error : public std::runtime_error {
 public:
  mcpm_dropin_error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

/* ---- Main_Functions.cpp operators ---------------------------------------- */
/* cell_update (cpp_headers.h:501; Main_Functions.cpp:2138-2749): bins the
   particles, rebuilds index[] and the cell lists, and sets each cell's
   weight, n, U_space, PIJ, M3, M4, M5, Q, DM2, DM4, Mp, T, Lambda and
   gamma_st, and for the FP models c, gamma, alpha, beta and phi.  T, rho
   and color are not used (the hybrid and SPH models). */
void cell_update(double *x1, double *x2, double *U1, double *U2, double *U3, struct GAS *gas,
                 struct CELLS *cells, struct BOX *box, int *index, double **T, double *rho,
                 int *color);
/* poisson_solver (cpp_headers.h:450; Main_Functions.cpp:3285-3581): the 1D
   screened-Poisson field, cells[].F2, and its kick U2 += F2 dt. */
void poisson_solver(double *U2, double *x1, double *x2, double *x3, struct GAS *gas,
                    struct BOX *box, struct CELLS *cells, int *index);
/* vlasov_taking_the_integral (cpp_headers.h:628; Main_Functions.cpp:
   3582-3942): the direct Sutherland integral, cells[].F2, and its kick. */
void vlasov_taking_the_integral(double *U2, struct GAS *gas, struct BOX *box,
                                struct CELLS *cells, int *index, double *x1, double *x2,
                                double *x3);
/* stream_particles (cpp_headers.h:495; Main_Functions.cpp:1494-1521) with
   the boundary of apply_BC for box->problem, in one pass: free flight of
   the slab (or plane) axes, then the walls; then the particles are binned
   again.  On the slab, x1 and x3 advance by U dt on the host when
   box->direction[0] or [2] is set.  x*_old are left to the caller. */
void stream_particles(double *U1, double *U2, double *U3, double *x1, double *x2, double *x3,
                      double *x1_old, double *x2_old, double *x3_old, struct GAS *gas,
                      struct BOX *box, struct CELLS *cells, int *index);
/* apply_BC (cpp_headers.h:498; Main_Functions.cpp:1451-1492): the walls are
   applied inside stream_particles; the reference's calls before and after
   streaming find every particle inside.  Here it only takes in particles
   the caller changed. */
void apply_BC(double *U1, double *U2, double *U3, double *x1, double *x2, double *x3,
              double *x1_old, double *x2_old, double *x3_old, struct GAS *gas, struct BOX *box,
              struct CELLS *cells, int *index, int *flag, double *xi_1, double *xi_2,
              double *xi_3, double *xi_1f, double *xi_2f, double *xi_3f, double *Mp1,
              double *Mp2, double *Mp3, double *T);
/* increase_collision_rate (cpp_headers.h:505; Main_Functions.cpp:
   1906-1915): the contact value Y(n). */
double increase_collision_rate(double n, struct GAS *gas);

/* ---- collisions.cpp ------------------------------------------------------ */
/* Collision_ESMC_new3 (cpp_headers.h:365; collisions.cpp:583-720) and
   Collision_ESMC_new4 (cpp_headers.h:364; collisions.cpp:722-837): the
   Enskog collisions of one step (box->step addresses the random stream),
   cells[].omega_max updated. */
void Collision_ESMC_new4(double *U1, double *U2, double *U3, double *x1, double *x2,
                         struct GAS *gas, struct BOX *box, struct CELLS *cells, int *index);
void Collision_ESMC_new3(double *U1, double *U2, double *U3, double *x1, double *x2,
                         struct GAS *gas, struct BOX *box, struct CELLS *cells, int *index);
/* Collision_DSMC_VHS (cpp_headers.h:612; collisions.cpp:2063-2141);
   cells[].crm updated.  The Collision scratch is not used. */
void Collision_DSMC_VHS(double *U1, double *U2, double *U3, double *x1, double *x2,
                        struct GAS *gas, struct BOX *box, struct CELLS *cells,
                        struct Collision *DSMC);

/* ---- Fokker_Planck.cpp --------------------------------------------------- */
/* random_generator (cpp_headers.h:362; Fokker_Planck.cpp:20-145): the FP
   noise comes from the library's counter-based stream keyed by box->step
   and is drawn inside Velocity_FP_parallel; the xi arrays are not written. */
void random_generator(double *xi_1, double *xi_2, double *xi_3, double *xi_1f, double *xi_2f,
                      double *xi_3f, struct CELLS *cells, struct BOX *box, struct GAS *gas,
                      int *index);
/* c_gamma_calculation (cpp_headers.h:467; Fokker_Planck.cpp:590-804) and
   alpha_calculation (cpp_headers.h:541; Fokker_Planck.cpp:200-587): the
   per-cell closures (both are computed together). */
void c_gamma_calculation(double *M1, double *M2, double *M3, struct GAS *gas, struct CELLS *cell,
                         struct BOX *box);
void alpha_calculation(double *M1, double *M2, double *M3, struct GAS *gas, struct CELLS *cells,
                       struct BOX *box);
/* Velocity_FP_parallel (cpp_headers.h:632; Fokker_Planck.cpp:1065-1182):
   the Langevin step of M1..M3 (the velocities) with the closures of the
   last cell_update; xi and Mp are not used (noise drawn on the device). */
void Velocity_FP_parallel(double *M1, double *M2, double *M3, struct GAS *gas, struct CELLS *cells,
                          struct BOX *box, int *index, double *xi_1, double *xi_2, double *xi_3,
                          double *Mp1, double *Mp2, double *Mp3);
/* DFP_extra_streaming (cpp_headers.h:406; Fokker_Planck.cpp:904-946): the
   FP_dense position drift is applied inside the next stream_particles. */
void DFP_extra_streaming(double *U1, double *U2, double *U3, double *x1, double *x2,
                         struct GAS *gas, struct BOX *box, struct CELLS *cells, int *index,
                         double dt);

/* ---- FEM.cpp (the plane, problem injection) ------------------------------ */
/* compute_SP_2D (cpp_headers.h:398; FEM.cpp:249-610) and compute_Vlasov_2D
   (cpp_headers.h:390; FEM.cpp:613-726): the nodal field, cells[].n_n, and
   its kick of U1 and U2. */
void compute_SP_2D(struct GAS *gas, struct BOX *box, struct CELLS *cells, double *x1, double *x2,
                   double *U1, double *U2, int *index);
void compute_Vlasov_2D(struct GAS *gas, struct BOX *box, struct CELLS *cells, double *x1,
                       double *x2, double *U1, double *U2, int *index);

/* ---- the library's own additions ----------------------------------------- */
/* Status of the last call (MCPM_OK or an MCPM_ERR_* code). */
int mcpm_dropin_status();
/* Releases the device state; the next call starts again from the caller's
   arrays. */
void mcpm_dropin_reset();

/* Layout of the reference's structs (g++ 11, x86-64). */
static_assert(sizeof(GAS) == 872, "GAS layout");
static_assert(sizeof(BOX) == 816, "BOX layout");
static_assert(sizeof(CELLS) == 1984, "CELLS layout");

#endif /* MCPM_DROPIN_H */
