// FP64 CPU restatement of the reference fluidLB D2Q9 time step, C++/OpenMP.
//
// TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline.  Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// it; the product never does.
//
// Same passes, same AoS layout pop[size][9] with ind = i * Ny_pad + j, and the
// same operation order as oracle/d2q9_oracle.py, which documents each pass
// against the reference files (pylabolt/solvers/fluidLB.py:206-253 and the
// kernels in pylabolt/parallel/cpu/).  Built with -ffp-contract=off so that
// no multiply-add is fused: the results equal the NumPy restatement bit for
// bit.
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {
const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
const int INV[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
}  // namespace

extern "C" {

// Boundary element kinds, in the order the Python side encodes them.
enum { OR_BB = 0, OR_VEL = 1, OR_PRES = 2, OR_ZG = 3, OR_PERIODIC = 4 };

typedef struct oracle_elem {
  int type;
  int n_nodes;
  const int64_t* nodes;
  int out[3];
  int inv[3];
  double vec[2];
  double scalar;
  int normal[2];
} oracle_elem;

typedef struct oracle_case {
  int nx_pad, ny_pad;
  const uint8_t* solid;
  const uint8_t* ghost;
  double* density;   // [size]
  double* velocity;  // [size][2]
  double* force;     // [size][2]
  double* pop;       // [size][9]
  double* pop_new;   // [size][9]
  double omega;
  double inv_cs_2, inv_cs_4, float_min;
  double w[9];
  int model;         // 0 BGK, 1 MRT
  int forcing;       // 0 none, 1 guo_linear, 2 guo_second_order
  int force_source;  // 0 gravity, 1 field
  double gravity[2];
  int x_periodic, y_periodic;
  int n_elements;
  const oracle_elem* elems;
} oracle_case;

static inline double feq_k(const oracle_case* c, int k, double rho, double ux,
                           double uy) {
  double u2 = ux * ux + uy * uy;
  double cu = (double)CX[k] * ux + (double)CY[k] * uy;
  return c->w[k] * rho *
         (((1.0 + c->inv_cs_2 * cu) + ((0.5 * c->inv_cs_4) * cu) * cu) -
          (0.5 * c->inv_cs_2) * u2);
}

static inline double guo_k(const oracle_case* c, int k, double ux, double uy,
                           double fx, double fy) {
  if (c->forcing == 1)
    return (c->w[k] * ((double)CX[k] * fx + (double)CY[k] * fy)) * c->inv_cs_2;
  double cu = (double)CX[k] * ux + (double)CY[k] * uy;
  double const_x = ((double)CX[k] - ux) * c->inv_cs_2 + (cu * (double)CX[k]) * c->inv_cs_4;
  double const_y = ((double)CY[k] - uy) * c->inv_cs_2 + (cu * (double)CY[k]) * c->inv_cs_4;
  return c->w[k] * (const_x * fx + const_y * fy);
}

// Build-defined MRT with Guo forcing and S = (1,1,1,1,1,1,1,omega,omega),
// reduced form of d2q9_oracle.mrt_collide, same operation order:
// q = feq (+ 0.5 Fbar), stress moments a7, a8 of f minus the closed-form
// stress of feq (+ half the closed-form stress of a second-order Fbar),
// r = ((1 - omega) * 0.25) a, f* = q +- r.
static void mrt_collide(const oracle_case* c, const double* pn, double* po, double rho,
                        double ux, double uy, double fx, double fy) {
  const double om = c->omega;
  const double w1 = c->w[1], w5 = c->w[5], ics4 = c->inv_cs_4;
  double q[9];
  for (int k = 0; k < 9; ++k) q[k] = feq_k(c, k, rho, ux, uy);
  double a7 = ((pn[1] + pn[3]) - (pn[2] + pn[4])) - ((w1 * ics4) * rho) * (ux * ux - uy * uy);
  double a8 = ((pn[5] + pn[7]) - (pn[6] + pn[8])) - (((4.0 * w5) * ics4) * rho) * (ux * uy);
  if (c->forcing != 0) {
    if (c->forcing == 2) {
      a7 = a7 + (w1 * ics4) * (ux * fx - uy * fy);
      a8 = a8 + ((2.0 * w5) * ics4) * (ux * fy + uy * fx);
    }
    for (int k = 0; k < 9; ++k) q[k] = q[k] + 0.5 * guo_k(c, k, ux, uy, fx, fy);
  }
  const double cs = (1.0 - om) * 0.25;
  const double r7 = cs * a7, r8 = cs * a8;
  po[0] = q[0];
  po[1] = q[1] + r7;
  po[2] = q[2] - r7;
  po[3] = q[3] + r7;
  po[4] = q[4] - r7;
  po[5] = q[5] + r8;
  po[6] = q[6] - r8;
  po[7] = q[7] + r8;
  po[8] = q[8] - r8;
}

int oracle_initialize(oracle_case* c) {
  const long size = (long)c->nx_pad * c->ny_pad;
#pragma omp parallel for schedule(static)
  for (long ind = 0; ind < size; ++ind) {
    if (c->solid[ind] || c->ghost[ind]) continue;
    double rho = c->density[ind];
    double ux = c->velocity[2 * ind], uy = c->velocity[2 * ind + 1];
    for (int k = 0; k < 9; ++k) {
      double f = feq_k(c, k, rho, ux, uy);
      c->pop[9 * ind + k] = f;
      c->pop_new[9 * ind + k] = f;
    }
  }
  return 0;
}

static void moments_and_collide(oracle_case* c) {
  const long size = (long)c->nx_pad * c->ny_pad;
  const double om = c->omega;
  // density, force, velocity: three passes as in the reference.
#pragma omp parallel for schedule(static)
  for (long ind = 0; ind < size; ++ind) {
    if (c->solid[ind] || c->ghost[ind]) continue;
    double s = 0.0;
    for (int k = 0; k < 9; ++k) s += c->pop_new[9 * ind + k];
    c->density[ind] = s;
  }
  if (c->force_source == 0) {
#pragma omp parallel for schedule(static)
    for (long ind = 0; ind < size; ++ind) {
      if (c->solid[ind] || c->ghost[ind]) continue;
      double rho = c->density[ind];
      c->force[2 * ind] = rho * c->gravity[0];
      c->force[2 * ind + 1] = rho * c->gravity[1];
    }
  }
#pragma omp parallel for schedule(static)
  for (long ind = 0; ind < size; ++ind) {
    if (c->solid[ind] || c->ghost[ind]) continue;
    double sx = 0.0, sy = 0.0;
    const double* pn = c->pop_new + 9 * ind;
    for (int k = 0; k < 9; ++k) {
      sx += (double)CX[k] * pn[k];
      sy += (double)CY[k] * pn[k];
    }
    double inv = 1.0 / (c->density[ind] + c->float_min);
    c->velocity[2 * ind] = sx * inv + (0.5 * c->force[2 * ind]) * inv;
    c->velocity[2 * ind + 1] = sy * inv + (0.5 * c->force[2 * ind + 1]) * inv;
  }
#pragma omp parallel for schedule(static)
  for (long ind = 0; ind < size; ++ind) {
    if (c->solid[ind] || c->ghost[ind]) continue;
    double rho = c->density[ind];
    double ux = c->velocity[2 * ind], uy = c->velocity[2 * ind + 1];
    double fx = c->force[2 * ind], fy = c->force[2 * ind + 1];
    const double* pn = c->pop_new + 9 * ind;
    double* po = c->pop + 9 * ind;
    if (c->model == 0) {
      for (int k = 0; k < 9; ++k) {
        double feq = feq_k(c, k, rho, ux, uy);
        double v = (1.0 - om) * pn[k] + om * feq;
        if (c->forcing != 0) v = v + (1.0 - 0.5 * om) * guo_k(c, k, ux, uy, fx, fy);
        po[k] = v;
      }
    } else {
      mrt_collide(c, pn, po, rho, ux, uy, fx, fy);
    }
  }
}

static void periodic_wrap(oracle_case* c) {
  const int nx = c->nx_pad, ny = c->ny_pad;
  if (c->x_periodic) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
      for (int k = 0; k < 9; ++k) {
        double l = c->pop[9 * ((long)(nx - 2) * ny + y) + k];
        double r = c->pop[9 * ((long)1 * ny + y) + k];
        c->pop[9 * ((long)0 * ny + y) + k] = l;
        c->pop[9 * ((long)(nx - 1) * ny + y) + k] = r;
      }
  }
  if (c->y_periodic) {
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
      for (int k = 0; k < 9; ++k) {
        double b = c->pop[9 * ((long)x * ny + ny - 2) + k];
        double t = c->pop[9 * ((long)x * ny + 1) + k];
        c->pop[9 * ((long)x * ny + 0) + k] = b;
        c->pop[9 * ((long)x * ny + ny - 1) + k] = t;
      }
  }
}

static void stream(oracle_case* c) {
  const long size = (long)c->nx_pad * c->ny_pad;
  const long ny = c->ny_pad;
#pragma omp parallel for schedule(static)
  for (long ind = 0; ind < size; ++ind) {
    if (c->solid[ind] || c->ghost[ind]) continue;
    long x = ind / ny, y = ind - x * ny;
    double rho = c->density[ind];
    c->pop_new[9 * ind] = c->pop[9 * ind];
    for (int k = 1; k < 9; ++k) {
      long nb = (x - CX[k]) * ny + (y - CY[k]);
      if (!c->solid[nb]) {
        c->pop_new[9 * ind + k] = c->pop[9 * nb + k];
      } else {
        int ki = INV[k];
        double temp = ((2.0 * c->w[ki]) * rho) * c->inv_cs_2 *
                      ((double)CX[ki] * c->velocity[2 * nb] +
                       (double)CY[ki] * c->velocity[2 * nb + 1]);
        c->pop_new[9 * ind + k] = c->pop[9 * ind + ki] - temp;
      }
    }
  }
}

static void boundaries(oracle_case* c, const double* streamed) {
  const long ny = c->ny_pad;
  for (int e = 0; e < c->n_elements; ++e) {
    const oracle_elem* el = &c->elems[e];
    if (el->type == OR_PERIODIC) continue;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < el->n_nodes; ++n) {
      long ind = el->nodes[n];
      if (c->solid[ind]) continue;
      if (el->type == OR_BB) {
        for (int q = 0; q < 3; ++q)
          c->pop_new[9 * ind + el->inv[q]] = c->pop[9 * ind + el->out[q]];
      } else if (el->type == OR_VEL) {
        double rho = c->density[ind];
        for (int q = 0; q < 3; ++q) {
          int o = el->out[q];
          double temp = ((2.0 * c->w[o]) * rho) * c->inv_cs_2 *
                        ((double)CX[o] * el->vec[0] + (double)CY[o] * el->vec[1]);
          c->pop_new[9 * ind + el->inv[q]] = c->pop[9 * ind + o] - temp;
        }
      } else if (el->type == OR_PRES) {
        long nb = ind + (long)el->normal[0] * ny + el->normal[1];
        double vx = c->velocity[2 * ind], vy = c->velocity[2 * ind + 1];
        double ex = vx + 0.5 * (vx - c->velocity[2 * nb]);
        double ey = vy + 0.5 * (vy - c->velocity[2 * nb + 1]);
        double u2 = ex * ex + ey * ey;
        for (int q = 0; q < 3; ++q) {
          int o = el->out[q];
          double cu = (double)CX[o] * ex + (double)CY[o] * ey;
          double temp = ((2.0 * c->w[o]) * el->scalar) *
                        (((1.0 + ((0.5 * c->inv_cs_4) * cu) * cu)) -
                         (0.5 * c->inv_cs_2) * u2);
          c->pop_new[9 * ind + el->inv[q]] = -c->pop[9 * ind + o] + temp;
        }
      } else if (el->type == OR_ZG) {
        long nb = ind + (long)el->normal[0] * ny + el->normal[1];
        for (int q = 0; q < 3; ++q) {
          int v = el->inv[q];
          c->pop_new[9 * ind + v] = streamed[9 * nb + v];
        }
      }
    }
  }
}

// Runs `steps` reference time steps with `n_threads` OpenMP threads (0 keeps
// the runtime's default).  Returns 0.
int oracle_run(oracle_case* c, int steps, int n_threads) {
  if (n_threads > 0) omp_set_num_threads(n_threads);
  const long size = (long)c->nx_pad * c->ny_pad;
  bool need_snapshot = false;
  for (int e = 0; e < c->n_elements; ++e)
    if (c->elems[e].type == OR_ZG) need_snapshot = true;
  std::vector<double> snap(need_snapshot ? (size_t)size * 9 : 0);
  for (int s = 0; s < steps; ++s) {
    moments_and_collide(c);
    periodic_wrap(c);
    stream(c);
    if (need_snapshot) std::memcpy(snap.data(), c->pop_new, sizeof(double) * size * 9);
    boundaries(c, need_snapshot ? snap.data() : nullptr);
  }
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
