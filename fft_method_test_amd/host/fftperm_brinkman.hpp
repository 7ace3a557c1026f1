// fftperm_brinkman.hpp -- the reference-side C++ binding of the Brinkman penalisation loop.
//
// common::fluid::solveBrinkman (src/brinkman/FluidBrinkman.cpp:282-507) builds its coefficient
// fields with helpers local to FluidBrinkman.cpp (resolvePhaseMask :67-131,
// buildCoefficientFields :152-191) and then iterates on the host.  The GPU branch goes right
// after those fields exist, inside solveBrinkman:
//
//     const CoefficientFields coeff = buildCoefficientFields(config, geometry, materialDb, phaseMask);
//     if (use_gpu)   // e.g. config.resources.backend != common::BackendType::CPU
//         return common::fluid::hip::solve_brinkman_fields(config, geometry, coeff.alpha, coeff.viscosity,
//                                                          coeff.porosity, coeff.meanFluidViscosity, force);
//
// so material lookup, phase resolution and every error message stay the reference's own, and
// computePermeabilityTensor (:509-553) needs no change.  Errors of the C ABI are rethrown as
// std::invalid_argument / std::runtime_error; a non-finite residual is the reference's
// "Brinkman solver became unstable" runtime_error (:473-477).
#pragma once

#include <array>
#include <chrono>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "fftperm.h"

#include "brinkman/FluidBrinkman.hpp"

namespace common::fluid::hip {

inline void throw_on_error(int rc) {
    if (rc == FFTPERM_OK) return;
    const std::string msg = fftperm_last_error();
    if (rc == FFTPERM_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// solveBrinkman's iteration on the GPU for one forcing vector.  alpha / viscosity are the
// per-voxel coefficient fields of buildCoefficientFields; porosity and meanFluidViscosity are
// copied into the result as the reference does (:289-291).  precision FFTPERM_PRECISION_MIXED
// runs the pressure-Poisson sweeps in FP32 (velocity, pressure and residual stay FP64).
inline SolveResult solve_brinkman_fields(const AppConfig& config, const Geometry& geometry,
                                         const std::vector<double>& alpha, const std::vector<double>& viscosity,
                                         double porosity, double meanFluidViscosity,
                                         const std::array<double, 3>& force,
                                         int precision = FFTPERM_PRECISION_FP64) {
    const std::size_t n = geometry.voxelCount();
    if (alpha.size() != n || viscosity.size() != n) {
        throw std::invalid_argument("solve_brinkman_fields: coefficient field size does not match the geometry");
    }
    SolveResult result;
    result.porosity = porosity;
    result.viscosity = meanFluidViscosity;
    std::vector<double> u(3 * n, 0.0);
    fftperm_brinkman_info info{};
    const auto start = std::chrono::steady_clock::now();
    throw_on_error(fftperm_brinkman_solve(geometry.nx, geometry.ny, geometry.nz, alpha.data(), viscosity.data(),
                                          config.fluid.relaxation, config.fluid.tolerance,
                                          config.fluid.maxIterations, force.data(), precision,
                                          result.superficialVelocity.data(), u.data(), &info));
    result.velocity.ux.assign(u.begin(), u.begin() + static_cast<std::ptrdiff_t>(n));
    result.velocity.uy.assign(u.begin() + static_cast<std::ptrdiff_t>(n), u.begin() + static_cast<std::ptrdiff_t>(2 * n));
    result.velocity.uz.assign(u.begin() + static_cast<std::ptrdiff_t>(2 * n), u.end());
    result.diagnostics.iterations = info.iterations;
    result.diagnostics.residual = info.residual;
    result.diagnostics.converged = info.converged != 0;
    result.diagnostics.elapsedSeconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    result.componentDiagnostics = {result.diagnostics, result.diagnostics, result.diagnostics};
    return result;
}

}  // namespace common::fluid::hip
