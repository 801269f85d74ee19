// Host-side helpers shared by the C ABI and the command-line driver.
#pragma once

#include <cstdint>

namespace g12io {

int save_checkpoint(const char* path, int Ns, int Nb, double t, int64_t steps, const double* G1,
                    const double* Gxx, const double* Gxy);
int load_checkpoint(const char* path, int Ns, int Nb, double* t, int64_t* steps, double* G1, double* Gxx,
                    double* Gxy);

// Hermitian eigen-decomposition (cyclic complex Jacobi, FP64).  a: n x n
// row-major interleaved complex, destroyed.  w: ascending eigenvalues;
// v: row-major n x n, column k is the eigenvector of w[k].
int hermitian_eigen(double* a, int n, double* w, double* v);

}  // namespace g12io
