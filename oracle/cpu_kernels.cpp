// OpenMP loops of the oracle's CPU baseline (test infrastructure: only the tests and the
// benchmark's CPU baseline load this library; the product path never does).
//
// The reference runs its stencils and sparse products as Numba prange loops
// (solver/numba_kernels.py:17-142 for the flux-form advection and diffusion, :347-354 for the CSR
// product).  These are the same loops in C++ with OpenMP over rows, evaluating every expression in
// the order oracle/ghost_oracle.py's NumPy code does, with FP contraction off, so the results equal
// the NumPy/SciPy ones bit for bit (tests/test_oracle_pins.py checks whole steps both ways).
#include <omp.h>
#include <cstdint>

extern "C" {

void gc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int gc_get_threads() { return omp_get_max_threads(); }

// y = A x for a CSR matrix, each row summed from 0.0 over its stored entries in order (SciPy's
// csr_matvec order).
void gc_csr_matvec(int64_t rows, const int32_t* indptr, const int32_t* indices, const double* data,
                   const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < rows; ++i) {
    double acc = 0.0;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) acc += data[k] * x[indices[k]];
    y[i] = acc;
  }
}

// Flux-form centred advection and diffusion of (u (ny, nx+1), v (ny+1, nx)) with unit apertures and
// no-slip walls: compute_flux_u/v + accumulate_rhs_u/v (numba_kernels.py:17-142), as
// ghost_oracle.flux_rhs.  adv is 1 with advection and 0 without.  work holds 2 (ny+1)(nx+1) doubles.
void gc_flux_rhs(int ny, int nx, const double* u, const double* v, double dx, double dy, double nu,
                 double adv, double rho, double* ru, double* rv, double* work) {
  const int nu1 = nx + 1;             // row length of u and of the corner arrays
  double* Fxy = work;                 // (ny+1, nx+1)
  double* Fyx = work + (int64_t)(ny + 1) * nu1;
  auto U = [&](int j, int i) { return u[(int64_t)j * nu1 + i]; };
  auto V = [&](int j, int i) { return v[(int64_t)j * nx + i]; };
  auto Fxx = [&](int j, int i) {      // cell-centre x flux of u (numba_kernels.py:23-35)
    const double uf = 0.5 * (U(j, i) + U(j, i + 1));
    return rho * nu * (U(j, i + 1) - U(j, i)) / dx * dy - adv * (rho * uf * uf * dy);
  };
  auto Fyy = [&](int j, int i) {      // cell-centre y flux of v (numba_kernels.py:82-94)
    const double vf = 0.5 * (V(j, i) + V(j + 1, i));
    return rho * nu * (V(j + 1, i) - V(j, i)) / dy * dx - adv * (rho * vf * vf * dx);
  };
#pragma omp parallel
  {
    // corner fluxes: F_xy of u (numba_kernels.py:37-62), F_yx of v (:96-132)
#pragma omp for schedule(static)
    for (int j = 0; j <= ny; ++j) {
      double* fxy = Fxy + (int64_t)j * nu1;
      double* fyx = Fyx + (int64_t)j * nu1;
      if (j == 0) {
        for (int i = 0; i <= nx; ++i) fxy[i] = 2.0 * rho * nu * U(0, i) / dy * dx;
      } else if (j == ny) {
        for (int i = 0; i <= nx; ++i) fxy[i] = -2.0 * rho * nu * U(ny - 1, i) / dy * dx;
      } else {
        for (int i = 0; i <= nx; ++i) {
          const double vface = i == 0 ? V(j, 0) : i == nx ? V(j, nx - 1) : 0.5 * (V(j, i - 1) + V(j, i));
          const double uadv = 0.5 * (U(j - 1, i) + U(j, i));
          fxy[i] = rho * nu * (U(j, i) - U(j - 1, i)) / dy * dx - adv * (rho * vface * uadv * dx);
        }
      }
      for (int i = 1; i < nx; ++i) {
        const double uface = j == 0 ? U(0, i) : j == ny ? U(ny - 1, i) : 0.5 * (U(j - 1, i) + U(j, i));
        const double vadv = 0.5 * (V(j, i - 1) + V(j, i));
        fyx[i] = rho * nu * (V(j, i) - V(j, i - 1)) / dx * dy - adv * (rho * uface * vadv * dy);
      }
      fyx[0] = 2.0 * rho * nu * V(j, 0) / dx * dy;
      const double uout = j == 0 ? U(0, nx) : j == ny ? U(ny - 1, nx) : 0.5 * (U(j - 1, nx) + U(j, nx));
      fyx[nx] = -rho * uout * V(j, nx - 1) * dy * adv;
    }
    // accumulate_rhs_u (numba_kernels.py:66-75)
#pragma omp for schedule(static)
    for (int j = 0; j < ny; ++j) {
      double* r = ru + (int64_t)j * nu1;
      const double* lo = Fxy + (int64_t)j * nu1;
      const double* hi = lo + nu1;
      r[0] = 0.0;
      double left = Fxx(j, 0);
      for (int i = 1; i < nx; ++i) {
        const double right = Fxx(j, i);
        r[i] = (right - left) + (hi[i] - lo[i]);
        left = right;
      }
      const double ext_x = -rho * U(j, nx) * U(j, nx) * dy * adv;
      r[nx] = ext_x - left + 0.5 * (hi[nx] - lo[nx]);
    }
    // accumulate_rhs_v (numba_kernels.py:135-142)
#pragma omp for schedule(static)
    for (int j = 0; j <= ny; ++j) {
      double* r = rv + (int64_t)j * nx;
      if (j == 0 || j == ny) {
        for (int i = 0; i < nx; ++i) r[i] = 0.0;
        continue;
      }
      const double* f = Fyx + (int64_t)j * nu1;
      for (int i = 0; i < nx; ++i) r[i] = (f[i + 1] - f[i]) + (Fyy(j, i) - Fyy(j - 1, i));
    }
  }
}

}  // extern "C"
