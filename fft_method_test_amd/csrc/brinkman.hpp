// brinkman.hpp -- GPU Brinkman penalisation loop (common::fluid::solveBrinkman).
#pragma once

#include <string>

namespace fftperm {

struct BrinkmanResult {
    int iterations = 0;
    double residual = 0.0;
    int converged = 0;
    std::string message;
};

// Returns 0 on success, 2 on a non-finite residual, 3 on a non-finite superficial velocity,
// 1 on a device error (message set).  alpha / viscosity / u are host arrays (u may be null).
int brinkman_run(int nx, int ny, int nz, const double* alpha, const double* viscosity, double relaxation, double tol,
                 int max_iterations, const double force[3], bool mixed, double U[3], double* u, BrinkmanResult* res);

}  // namespace fftperm
