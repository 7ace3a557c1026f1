// ref_harness.cpp -- extern "C" entry points over the reference's own C++ classes.
//
// TEST INFRASTRUCTURE ONLY.  build_ref.sh compiles this file together with the
// reference sources where they lie (src/geometry, src/bignonnet, src/solver,
// src/post, src/brinkman, src/fft/FFTWBackend.cpp, ...) into oracle/_ref/libref.so.
// FFTW3 is not installed in this image, so FFTWBackend.cpp is compiled without
// PERMEABILITY_HAVE_FFTW3*; it then takes its own fallback branches
// (src/fft/FFTWBackend.cpp:309-315 forward = real -> complex copy,
// :374-380 inverse = real part, no 1/N).  GreenOperator::apply therefore
// multiplies by the reference's Green tensor pointwise, a symmetric operator:
// everything above the FFT (support, Green tensor, f0, RHS, operator, MINRES,
// recovery, K assembly, Brinkman) runs as the reference's own compiled code, and
// the oracle's identity-transform mode is compared against it.
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "bignonnet/BackgroundForce.hpp"
#include "bignonnet/ForceSupport.hpp"
#include "bignonnet/GreenDiscretization.hpp"
#include "bignonnet/GreenOperator.hpp"
#include "bignonnet/TrialForceField.hpp"
#include "bignonnet/VariationalOperator.hpp"
#include "bignonnet/VariationalRHS.hpp"
#include "bignonnet/VelocityRecovery.hpp"
#include "brinkman/FluidBrinkman.hpp"
#include "fft/FFTWBackend.hpp"
#include "geometry/BinaryMedium.hpp"
#include "geometry/InterfaceVoxels.hpp"
#include "io/VoxelIO.hpp"
#include "materials/MaterialDatabase.hpp"
#include "post/Diagnostics.hpp"
#include "post/PermeabilityTensor.hpp"
#include "solver/MinresSolver.hpp"

using namespace permeability;

namespace {
thread_local std::string g_err;

struct RefCtx {
    Grid3D grid;
    BinaryMedium medium;
    ForceSupport support;
    FFTWBackend fft;
    GreenOperator green;
    BackgroundForce f0;
    VariationalRHS rhs;
    VariationalOperator op;
    VelocityRecovery recovery;
    double mu;

    RefCtx(std::size_t nx, std::size_t ny, std::size_t nz, double h, const unsigned char* mask,
           int support_mode, double mu_, int p_radius)
        : grid(nx, ny, nz, h, h, h),
          medium(grid, std::vector<unsigned char>(mask, mask + nx * ny * nz)),
          support(medium, support_mode == 1 ? SupportMode::InterfaceOnly : SupportMode::FullSolid),
          fft(),
          green(grid, GreenDiscretization{}, mu_, fft, 1, p_radius),
          f0(medium, support),
          rhs(support, green, f0),
          op(support, green),
          recovery(support, green, f0),
          mu(mu_) {}
};
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_create(std::size_t nx, std::size_t ny, std::size_t nz, double h, const unsigned char* mask,
                 int support_mode, double mu, int p_radius) {
    try {
        return new RefCtx(nx, ny, nz, h, mask, support_mode, mu, p_radius);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_destroy(void* p) { delete static_cast<RefCtx*>(p); }

std::uint64_t ref_num_active(void* p) { return static_cast<RefCtx*>(p)->support.num_active_voxels(); }

void ref_get_support(void* p, std::uint64_t* active, std::int64_t* v2l) {
    auto* c = static_cast<RefCtx*>(p);
    const auto& a = c->support.active_voxels();
    for (std::size_t i = 0; i < a.size(); ++i) active[i] = a[i];
    if (v2l) {
        for (std::size_t i = 0; i < c->grid.total_size(); ++i) {
            v2l[i] = c->support.contains_voxel(i) ? static_cast<std::int64_t>(c->support.local_index(i)) : -1;
        }
    }
}

void ref_green_tensor(void* p, double* t6) {
    auto* c = static_cast<RefCtx*>(p);
    const std::size_t n = c->grid.total_size();
    for (std::size_t i = 0; i < n; ++i) {
        const auto t = c->green.tensor_components(i);
        for (int g = 0; g < 6; ++g) t6[g * n + i] = t[g];
    }
}

int ref_green_has_nan(void* p) { return static_cast<RefCtx*>(p)->green.has_nan_in_tensor() ? 1 : 0; }

void ref_background_force(void* p, const double grad[3], double* out) {
    auto* c = static_cast<RefCtx*>(p);
    const VectorField3D f = c->f0.build(Real3{grad[0], grad[1], grad[2]});
    const std::size_t n = c->grid.total_size();
    for (std::size_t i = 0; i < n; ++i) {
        out[i] = f.x()[i];
        out[n + i] = f.y()[i];
        out[2 * n + i] = f.z()[i];
    }
}

void ref_green_apply(void* p, const double* fin, double* uout) {
    auto* c = static_cast<RefCtx*>(p);
    const std::size_t n = c->grid.total_size();
    VectorField3D f(c->grid), u(c->grid);
    for (std::size_t i = 0; i < n; ++i) {
        f.x()[i] = fin[i];
        f.y()[i] = fin[n + i];
        f.z()[i] = fin[2 * n + i];
    }
    c->green.apply(f, u);
    for (std::size_t i = 0; i < n; ++i) {
        uout[i] = u.x()[i];
        uout[n + i] = u.y()[i];
        uout[2 * n + i] = u.z()[i];
    }
}

void ref_rhs(void* p, const double grad[3], double* b) {
    auto* c = static_cast<RefCtx*>(p);
    const std::vector<Scalar> v = c->rhs.build(Real3{grad[0], grad[1], grad[2]});
    std::copy(v.begin(), v.end(), b);
}

void ref_operator_apply(void* p, const double* x, double* y) {
    auto* c = static_cast<RefCtx*>(p);
    const std::size_t L = c->op.size();
    std::vector<Scalar> in(x, x + L), out;
    c->op.apply(in, out);
    std::copy(out.begin(), out.end(), y);
}

int ref_minres(void* p, const double* b, const double* x0, double tol, std::uint64_t maxit,
               double* x, double* history, std::uint64_t* hist_len, std::uint64_t* iters,
               double* final_rel, int* converged) {
    auto* c = static_cast<RefCtx*>(p);
    try {
        const std::size_t L = c->op.size();
        MinresSolver solver(tol, maxit);
        std::vector<Scalar> rhs(b, b + L);
        std::vector<Scalar> guess;
        if (x0) guess.assign(x0, x0 + L);
        const SolverResult r = solver.solve(c->op, rhs, guess);
        std::copy(r.solution.begin(), r.solution.end(), x);
        if (history) std::copy(r.residual_history.begin(), r.residual_history.end(), history);
        *hist_len = r.residual_history.size();
        *iters = r.iterations;
        *final_rel = r.final_rel_residual;
        *converged = r.converged ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void ref_recover(void* p, const double grad[3], const double* x, double* u, double avg[3],
                 double avg_support[3]) {
    auto* c = static_cast<RefCtx*>(p);
    const std::size_t L = c->op.size();
    const std::vector<Scalar> xs(x, x + L);
    const Real3 g{grad[0], grad[1], grad[2]};
    const VectorField3D v = c->recovery.recover(g, xs);
    const Real3 a = c->recovery.average_over_domain(v);
    const Real3 s = c->recovery.average_over_support(v);
    for (int i = 0; i < 3; ++i) {
        avg[i] = a[i];
        avg_support[i] = s[i];
    }
    if (u) {
        const std::size_t n = c->grid.total_size();
        for (std::size_t i = 0; i < n; ++i) {
            u[i] = v.x()[i];
            u[n + i] = v.y()[i];
            u[2 * n + i] = v.z()[i];
        }
    }
}

// PermeabilityRunner::run K assembly (src/app/PermeabilityRunner.cpp:757-777) through the
// reference's PermeabilityTensor::set_column.
void ref_set_column_k(double mu, int col, const double avg[3], double K[9]) {
    PermeabilityTensor k;
    const Real3 colK{-mu * avg[0], -mu * avg[1], -mu * avg[2]};
    k.set_column(static_cast<std::size_t>(col), colK);
    for (int r = 0; r < 3; ++r) K[r * 3 + col] = k.data()[r * 3 + col];
}

double ref_residual_norm(void* p, const double* x, const double* b) {
    auto* c = static_cast<RefCtx*>(p);
    const std::size_t L = c->op.size();
    return Diagnostics::residual_norm(c->op, std::vector<Scalar>(x, x + L), std::vector<Scalar>(b, b + L));
}

double ref_symmetry_gap(void* p, const double* x, const double* y) {
    auto* c = static_cast<RefCtx*>(p);
    const std::size_t L = c->op.size();
    return Diagnostics::relative_symmetry_gap(c->op, std::vector<Scalar>(x, x + L),
                                              std::vector<Scalar>(y, y + L));
}

// common::fluid::solveBrinkman with an auto-built material database (no phase tags), as
// PermeabilityRunner's run_brinkman does without a material file (src/app/PermeabilityRunner.cpp:261-281).
int ref_brinkman_solve(std::size_t nx, std::size_t ny, std::size_t nz, const std::uint32_t* ids,
                       std::uint32_t fluid_id, double viscosity, double penalty, double relaxation,
                       double tol, int maxit, const double force[3], double U[3], double* u,
                       int* iters, double* residual, int* converged, double* porosity) {
    try {
        common::Geometry geo;
        geo.nx = nx;
        geo.ny = ny;
        geo.nz = nz;
        geo.materialIds.assign(ids, ids + nx * ny * nz);
        common::MaterialDatabase db;
        common::MaterialProperties props;
        props.values["viscosity"] = viscosity;
        db.insert(fluid_id, std::move(props));
        common::AppConfig cfg;
        cfg.fluid.fluidMaterialId = fluid_id;
        cfg.fluid.solidPenalty = penalty;
        cfg.fluid.relaxation = relaxation;
        cfg.fluid.tolerance = tol;
        cfg.fluid.maxIterations = maxit;
        cfg.fluid.requireSingleSolidMaterialId = false;
        const std::array<double, 3> f{force[0], force[1], force[2]};
        const common::fluid::SolveResult r = common::fluid::solveBrinkman(cfg, geo, db, f);
        for (int i = 0; i < 3; ++i) U[i] = r.superficialVelocity[i];
        if (u) {
            const std::size_t n = nx * ny * nz;
            std::copy(r.velocity.ux.begin(), r.velocity.ux.end(), u);
            std::copy(r.velocity.uy.begin(), r.velocity.uy.end(), u + n);
            std::copy(r.velocity.uz.begin(), r.velocity.uz.end(), u + 2 * n);
        }
        *iters = r.diagnostics.iterations;
        *residual = r.diagnostics.residual;
        *converged = r.diagnostics.converged ? 1 : 0;
        *porosity = r.porosity;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int ref_brinkman_tensor(std::size_t nx, std::size_t ny, std::size_t nz, const std::uint32_t* ids,
                        std::uint32_t fluid_id, double viscosity, double penalty, double relaxation,
                        double tol, int maxit, double forcing, double voxel_size, double K[9],
                        int iters[3], double residual[3]) {
    try {
        common::Geometry geo;
        geo.nx = nx;
        geo.ny = ny;
        geo.nz = nz;
        geo.materialIds.assign(ids, ids + nx * ny * nz);
        common::MaterialDatabase db;
        common::MaterialProperties props;
        props.values["viscosity"] = viscosity;
        db.insert(fluid_id, std::move(props));
        common::AppConfig cfg;
        cfg.fluid.fluidMaterialId = fluid_id;
        cfg.fluid.solidPenalty = penalty;
        cfg.fluid.relaxation = relaxation;
        cfg.fluid.tolerance = tol;
        cfg.fluid.maxIterations = maxit;
        cfg.fluid.forcingMagnitude = forcing;
        cfg.fluid.voxelSize = voxel_size;
        cfg.fluid.requireSingleSolidMaterialId = false;
        const common::fluid::PermeabilityResult r = common::fluid::computePermeabilityTensor(cfg, geo, db);
        for (int i = 0; i < 9; ++i) K[i] = r.tensor[i];
        for (int i = 0; i < 3; ++i) {
            iters[i] = r.loadDiagnostics[i].iterations;
            residual[i] = r.loadDiagnostics[i].residual;
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// VoxelIO::load (mask) and VoxelIO::load_geometry (raw ids), src/io/VoxelIO.cpp:48-252.
int ref_load_mask(const char* path, int fluid_id, std::uint64_t dims[3], unsigned char* mask,
                  std::uint64_t cap) {
    try {
        const BinaryMedium m = VoxelIO::load(path, 1.0, fluid_id);
        dims[0] = m.grid().nx();
        dims[1] = m.grid().ny();
        dims[2] = m.grid().nz();
        if (mask) {
            if (cap < m.mask().size()) {
                g_err = "capacity";
                return -2;
            }
            std::copy(m.mask().begin(), m.mask().end(), mask);
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int ref_load_geometry(const char* path, std::uint64_t dims[3], std::uint32_t* ids, std::uint64_t cap) {
    try {
        const common::Geometry g = VoxelIO::load_geometry(path);
        dims[0] = g.nx;
        dims[1] = g.ny;
        dims[2] = g.nz;
        if (ids) {
            if (cap < g.materialIds.size()) {
                g_err = "capacity";
                return -2;
            }
            std::copy(g.materialIds.begin(), g.materialIds.end(), ids);
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
