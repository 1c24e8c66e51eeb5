// Drop-in check of the C++ API: spla::pgemm_sbs and spla::pgemm_ssbtr with
// std::complex<double> through spla::Context / spla::MatrixDistribution, and
// the exception on an invalid argument. Prints "max error E" and "exception
// ok"; returns 0 on success.
#include <mpi.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "spla/pgemm_ssbtr.hpp"
#include "spla/spla.hpp"

using cd = std::complex<double>;

int main() {
  int rc = 0;
  try {
    const int mLocal = 300, n = 70, k = 90;
    std::vector<cd> A(static_cast<std::size_t>(mLocal) * k), B(static_cast<std::size_t>(k) * n),
        C(static_cast<std::size_t>(mLocal) * n, cd(1.0, -1.0));
    for (std::size_t i = 0; i < A.size(); ++i) A[i] = cd(std::sin(0.1 * i), std::cos(0.3 * i));
    for (std::size_t i = 0; i < B.size(); ++i) B[i] = cd(std::cos(0.2 * i), std::sin(0.7 * i));
    const std::vector<cd> C0 = C;
    spla::Context ctx(SPLA_PU_GPU);
    auto dist = spla::MatrixDistribution::create_blacs_block_cyclic(MPI_COMM_SELF, 'C', 1, 1, 32, 32);
    const cd alpha(0.5, 0.25), beta(2.0, 0.0);
    spla::pgemm_sbs(mLocal, n, k, alpha, A.data(), mLocal, B.data(), k, 0, 0, dist, beta, C.data(),
                    mLocal, ctx);
    double maxErr = 0.0, maxVal = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < mLocal; ++i) {
        cd s(0.0, 0.0);
        for (int l = 0; l < k; ++l) s += A[i + static_cast<std::size_t>(l) * mLocal] * B[l + static_cast<std::size_t>(j) * k];
        const cd ref = alpha * s + beta * C0[i + static_cast<std::size_t>(j) * mLocal];
        maxErr = std::max(maxErr, std::abs(C[i + static_cast<std::size_t>(j) * mLocal] - ref));
        maxVal = std::max(maxVal, std::abs(ref));
      }
    std::printf("max error %.3e\n", maxErr / maxVal);
    if (maxErr / maxVal > 1e-12) rc = 1;

    // ssb requires a (conjugate) transpose of A: SPLA throws InvalidParameterError
    try {
      spla::pgemm_ssbtr(4, 4, 4, SPLA_OP_NONE, cd(1.0), A.data(), 4, B.data(), 4, cd(0.0),
                        C.data(), 4, 0, 0, SPLA_FILL_MODE_UPPER, dist, ctx);
      rc = 1;
    } catch (const spla::InvalidParameterError& e) {
      std::printf("exception ok: %s (code %d)\n", e.what(), static_cast<int>(e.error_code()));
    }
  } catch (const spla::GenericError& e) {
    std::fprintf(stderr, "SPLA error %d: %s\n", static_cast<int>(e.error_code()), e.what());
    return 2;
  }
  return rc;
}
