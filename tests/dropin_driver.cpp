// dropin_driver.cpp — test driver: the reference's time loop (main.cpp:
// 389-669, the ESMC slab branch) run through the drop-in C++ symbols of
// libmcpm_dropin.so, on structs and arrays set up the way the reference's
// own code sets them up.
//
//   dropin_driver <in> <out> <steps> <operator>
//
// <in>: int64 N, int32 n_cells, int32 problem, then doubles length, sigma,
// m, kb, b, Fn, cell volume, dt, T_wall_1, T_wall_2, U_wall_1, U_wall_2,
// mu, T0, visp, then N doubles each of x2, U1, U2, U3 and n_cells doubles
// of omega_max.
// <operator>: 3 or 4 (ESMC, Collision_ESMC_new3 / _new4), 5 (FP_dense:
// random_generator, Velocity_FP_parallel and DFP_extra_streaming,
// main.cpp:415-470).
// <out>: the cells' counts (int32), omega_max and n (doubles), then the
// particles (x2, U1, U2, U3, doubles) listed cell by cell through
// cells[c].indices_inside, and index[] (int32) in the caller's order.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "mcpm_dropin.h"

template <typename T>
static void rd(FILE *f, T *v, size_t n) {
  if (fread(v, sizeof(T), n, f) != n) {
    fprintf(stderr, "short input\n");
    exit(2);
  }
}

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: dropin_driver in out steps operator\n");
    return 2;
  }
  const int steps = atoi(argv[3]);
  const int op = atoi(argv[4]);
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t N;
  int32_t nc, problem;
  rd(f, &N, 1);
  rd(f, &nc, 1);
  rd(f, &problem, 1);
  double v[15];
  rd(f, v, 15);
  std::vector<double> x1(N, 0.0), x2(N), x3(N, 0.0), U1(N), U2(N), U3(N);
  rd(f, x2.data(), N);
  rd(f, U1.data(), N);
  rd(f, U2.data(), N);
  rd(f, U3.data(), N);
  std::vector<double> om(nc);
  rd(f, om.data(), nc);
  fclose(f);

  // the structs as readSettingsFile / make_cells leave them (myIO.cpp)
  GAS gas{};
  BOX box{};
  gas.model = op == 5 ? "FP_dense" : "ESMC";
  gas.mu = v[12];
  gas.T0 = v[13];
  gas.visp = v[14];
  gas.N = N;
  gas.sigma = v[1];
  gas.m = v[2];
  gas.kb = v[3];
  gas.b = v[4];
  gas.Fn = v[5];
  gas.delta_t = v[7];
  gas.long_range = 0;
  box.problem = problem;
  box.N[0] = 1;
  box.N[1] = nc;
  box.N[2] = 1;
  box.Len[0] = 1.0;
  box.Len[1] = v[0];
  box.Len[2] = 1.0;
  box.delta_dim[1] = v[0] / nc;
  box.T_wall_1 = v[8];
  box.T_wall_2 = v[9];
  box.U_wall_1 = v[10];
  box.U_wall_2 = v[11];
  box.direction[0] = 0;
  box.direction[1] = 1;
  box.direction[2] = 0;
  std::vector<CELLS> cells(nc);
  for (int c = 0; c < nc; ++c) {
    memset(&cells[c], 0, sizeof(CELLS));
    cells[c].volume = v[6];
    cells[c].n_ratio = 1;
    cells[c].omega_max = om[c];
  }
  std::vector<int> index(N, 0), flag(N, 0), color(N, 0);
  std::vector<double> x1o(N), x2o(N), x3o(N), rho(N, 0.0), Tp(N, 0.0), xi(N, 0.0);
  double *T = Tp.data();
  try {
    for (int step = 0; step < steps; ++step) {
      box.step = step;
      cell_update(x1.data(), x2.data(), U1.data(), U2.data(), U3.data(), &gas, cells.data(), &box,
                  index.data(), &T, rho.data(), color.data());
      if (op == 5) {
        // random_generator, then velocity_update's FP_dense branch
        // (Main_Functions.cpp:1566-1570) and the extra streaming of FP_dense
        random_generator(xi.data(), xi.data(), xi.data(), xi.data(), xi.data(), xi.data(),
                         cells.data(), &box, &gas, index.data());
        Velocity_FP_parallel(U1.data(), U2.data(), U3.data(), &gas, cells.data(), &box,
                             index.data(), xi.data(), xi.data(), xi.data(), xi.data(), xi.data(),
                             xi.data());
        DFP_extra_streaming(U1.data(), U2.data(), U3.data(), x1.data(), x2.data(), &gas, &box,
                            cells.data(), index.data(), gas.delta_t);
      } else if (op == 4) {
        // velocity_update, ESMC branch (Main_Functions.cpp:1555-1561)
        Collision_ESMC_new4(U1.data(), U2.data(), U3.data(), x1.data(), x2.data(), &gas, &box,
                            cells.data(), index.data());
      } else {
        Collision_ESMC_new3(U1.data(), U2.data(), U3.data(), x1.data(), x2.data(), &gas, &box,
                            cells.data(), index.data());
      }
      apply_BC(U1.data(), U2.data(), U3.data(), x1.data(), x2.data(), x3.data(), x1o.data(),
               x2o.data(), x3o.data(), &gas, &box, cells.data(), index.data(), flag.data(),
               xi.data(), xi.data(), xi.data(), xi.data(), xi.data(), xi.data(), xi.data(),
               xi.data(), xi.data(), T);
      x1o = x1;  // store_old_positions (Main_Functions.cpp:5736-5759)
      x2o = x2;
      x3o = x3;
      stream_particles(U1.data(), U2.data(), U3.data(), x1.data(), x2.data(), x3.data(),
                       x1o.data(), x2o.data(), x3o.data(), &gas, &box, cells.data(),
                       index.data());
      apply_BC(U1.data(), U2.data(), U3.data(), x1.data(), x2.data(), x3.data(), x1o.data(),
               x2o.data(), x3o.data(), &gas, &box, cells.data(), index.data(), flag.data(),
               xi.data(), xi.data(), xi.data(), xi.data(), xi.data(), xi.data(), xi.data(),
               xi.data(), xi.data(), T);
    }
    // the next loop iteration's binning
    cell_update(x1.data(), x2.data(), U1.data(), U2.data(), U3.data(), &gas, cells.data(), &box,
                index.data(), &T, rho.data(), color.data());
  } catch (const mcpm_dropin_error &e) {
    fprintf(stderr, "drop-in error %d: %s\n", e.code(), e.what());
    return 1;
  }
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  for (int c = 0; c < nc; ++c) fwrite(&cells[c].num_inside, sizeof(int), 1, o);
  for (int c = 0; c < nc; ++c) fwrite(&cells[c].omega_max, sizeof(double), 1, o);
  for (int c = 0; c < nc; ++c) fwrite(&cells[c].n, sizeof(double), 1, o);
  const double *arr[4] = {x2.data(), U1.data(), U2.data(), U3.data()};
  for (int q = 0; q < 4; ++q)
    for (int c = 0; c < nc; ++c)
      for (int u = 0; u < cells[c].num_inside; ++u)
        fwrite(&arr[q][cells[c].indices_inside[u]], sizeof(double), 1, o);
  fwrite(index.data(), sizeof(int), N, o);
  fclose(o);
  for (int c = 0; c < nc; ++c) free(cells[c].indices_inside);
  mcpm_dropin_reset();
  return 0;
}
