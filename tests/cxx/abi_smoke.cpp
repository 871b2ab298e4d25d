// A C++ caller of the C ABI (include/mpdg_heat.h) as a MeltPoolDG-side adapter would be:
// creates an FP64 Q2 operator on 5 x 4 x 3 cells, applies it to the constant vector and checks
// sum_i (J 1)_i = c_m |Omega| (the stiffness part annihilates constants; c_m = rho c_p / dt),
// then checks that vmult and vmult_noalloc agree and that errors come back as codes.
// Prints "abi_smoke ok" and returns 0 on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/mpdg_heat.h"

#define CHECK(x)                                                          \
  do {                                                                    \
    const int rc_ = (x);                                                  \
    if (rc_ != MPDG_OK) {                                                 \
      std::fprintf(stderr, "%s failed: %d (line %d)\n", #x, rc_, __LINE__); \
      return 1;                                                           \
    }                                                                     \
  } while (0)

int main() {
  mpdg_heat_desc d{};
  d.degree = 2;
  d.n_cells[0] = 5;
  d.n_cells[1] = 4;
  d.n_cells[2] = 3;
  d.extent[0] = 1.0;
  d.extent[1] = 0.8;
  d.extent[2] = 0.6;
  d.precision = MPDG_FP64;
  d.dirichlet_faces = 0;
  d.rho_cp = 2.0;
  mpdg_heat_op *op = nullptr;
  CHECK(mpdg_heat_op_create(&d, &op));
  const double dt = 1e-3;
  CHECK(mpdg_heat_op_reset_time_increment(op, dt));
  CHECK(mpdg_heat_op_set_conductivity_constant(op, 3.0));
  const long n = mpdg_heat_op_n_dofs(op);
  if (n != 11L * 9L * 7L) {
    std::fprintf(stderr, "n_dofs %ld\n", n);
    return 1;
  }
  double *x = nullptr, *y = nullptr, *z = nullptr;
  if (hipMalloc(&x, n * sizeof(double)) != hipSuccess || hipMalloc(&y, n * sizeof(double)) != hipSuccess ||
      hipMalloc(&z, n * sizeof(double)) != hipSuccess)
    return 1;
  std::vector<double> h(n, 1.0);
  if (hipMemcpy(x, h.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 1;
  CHECK(mpdg_heat_op_vmult(op, y, x, nullptr));
  CHECK(mpdg_heat_op_vmult_noalloc(op, z, x, nullptr));
  std::vector<double> hy(n), hz(n);
  if (hipMemcpy(hy.data(), y, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  if (hipMemcpy(hz.data(), z, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  double sum = 0.0;
  for (long i = 0; i < n; ++i) {
    sum += hy[i];
    if (hy[i] != hz[i]) {
      std::fprintf(stderr, "vmult and vmult_noalloc differ at %ld\n", i);
      return 1;
    }
  }
  const double expect = d.rho_cp / dt * d.extent[0] * d.extent[1] * d.extent[2];
  if (std::fabs(sum - expect) > 1e-10 * expect) {
    std::fprintf(stderr, "sum %.17g expected %.17g\n", sum, expect);
    return 1;
  }
  // errors are codes, not exceptions
  if (mpdg_heat_op_reset_time_increment(op, 0.0) != MPDG_ERR_ZERO_TIME_INCREMENT) return 1;
  if (mpdg_heat_op_vmult(nullptr, y, x, nullptr) != MPDG_ERR_INVALID) return 1;
  CHECK(mpdg_heat_op_destroy(op));
  (void)hipFree(x);
  (void)hipFree(y);
  (void)hipFree(z);
  std::printf("abi_smoke ok: sum(J 1) = %.15g\n", sum);
  return 0;
}
