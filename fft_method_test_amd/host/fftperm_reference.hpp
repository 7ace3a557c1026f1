// fftperm_reference.hpp -- the reference-side C++ binding of libfftperm.so.
//
// Drop this header into lcolabella/FFT_Method_Test (compile with `-I src -I include` as the
// reference does, plus the directory of fftperm.h) and link -lfftperm.  It implements the
// reference's own plug points over the C ABI of include/fftperm.h:
//
//   permeability::hip::HipFFTBackend          : IFFTBackend      (src/fft/IFFTBackend.hpp:12-50)
//   permeability::hip::HipVariationalOperator : ILinearOperator  (src/solver/ILinearOperator.hpp:10-16)
//   permeability::hip::HipMinresSolver        MinresSolver::solve's signature on the device
//                                             (src/solver/MinresSolver.hpp:16-25)
//   permeability::hip::solve_one_direction    PermeabilityRunner's solve_one_direction
//                                             (src/app/PermeabilityRunner.cpp:173-242)
//   permeability::hip::solve_permeability_tensor   the gradient=all K assembly
//                                             (src/app/PermeabilityRunner.cpp:757-777,
//                                              src/post/PermeabilityTensor.cpp:11-18)
//   permeability::hip::make_gpu_fft_backend   the ComputeBackend::GPU case of create_fft_backend
//                                             (src/fft/FFTBackendFactory.cpp:13-34)
//
// Scalar is float in the reference's default build and double with
// -DFFT_METHOD_PRECISION_DOUBLE (src/core/Types.hpp:14-21).  The library computes in FP64 (or in
// its MIXED mode); with Scalar = float every call widens its inputs to double and rounds its
// outputs back to float at the boundary, with Scalar = double the caller's buffers are passed
// through without a copy.
//
// Errors: every non-zero status of the C ABI is rethrown as the exception type the reference
// uses for that failure (std::invalid_argument, std::out_of_range, std::runtime_error), with
// fftperm_last_error() as the message.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "fftperm.h"

#include "core/Field3D.hpp"
#include "core/Grid3D.hpp"
#include "core/Types.hpp"
#include "fft/IFFTBackend.hpp"
#include "geometry/BinaryMedium.hpp"
#include "post/PermeabilityTensor.hpp"
#include "solver/ILinearOperator.hpp"
#include "solver/SolverResult.hpp"

namespace permeability::hip {

// Status code of the C ABI -> the reference's exception types.
inline void throw_on_error(int rc) {
    if (rc == FFTPERM_OK) return;
    const std::string msg = fftperm_last_error();
    switch (rc) {
        case FFTPERM_EINVAL:
            throw std::invalid_argument(msg);
        case FFTPERM_ERANGE:
            throw std::out_of_range(msg);
        default:
            throw std::runtime_error(msg);
    }
}

// Read-only double view of a Scalar vector: the vector itself when Scalar is double, else a
// widened copy held by the view.
template <class T>
class DoubleInT {
public:
    explicit DoubleInT(const std::vector<T>& v) {
        if constexpr (std::is_same_v<T, double>) {
            ptr_ = v.data();
        } else {
            copy_.assign(v.begin(), v.end());
            ptr_ = copy_.data();
        }
    }
    [[nodiscard]] const double* data() const noexcept { return ptr_; }

private:
    std::vector<double> copy_;
    const double* ptr_ = nullptr;
};

// Writable double buffer of n entries that lands in a Scalar vector: writes in place when Scalar
// is double, else through a double buffer rounded into the vector by commit().
template <class T>
class DoubleOutT {
public:
    DoubleOutT(std::vector<T>& v, std::size_t n) : dst_(v) {
        dst_.resize(n);
        if constexpr (std::is_same_v<T, double>) {
            ptr_ = dst_.data();
        } else {
            buf_.assign(n, 0.0);
            ptr_ = buf_.data();
        }
    }
    [[nodiscard]] double* data() noexcept { return ptr_; }
    void commit() {
        if constexpr (!std::is_same_v<T, double>) {
            for (std::size_t i = 0; i < buf_.size(); ++i) dst_[i] = static_cast<T>(buf_[i]);
        }
    }

private:
    std::vector<T>& dst_;
    std::vector<double> buf_;
    double* ptr_ = nullptr;
};

using DoubleIn = DoubleInT<Scalar>;
using DoubleOut = DoubleOutT<Scalar>;

inline fftperm_grid to_fftperm_grid(const Grid3D& g) {
    return fftperm_grid{g.nx(), g.ny(), g.nz(),
                        static_cast<double>(g.hx()), static_cast<double>(g.hy()), static_cast<double>(g.hz()),
                        static_cast<double>(g.lx()), static_cast<double>(g.ly()), static_cast<double>(g.lz())};
}

// One library context: the grid's FFT plans and Green tensor, the medium's support map and the
// device vectors of the solves, resident on the GPU for the context's lifetime.  Plays the part
// of the reference's per-load-case GreenOperator + ForceSupport pair, built once for all cases.
class Context {
public:
    Context(const Grid3D& grid, Scalar viscosity, int p_radius, int precision = FFTPERM_PRECISION_FP64)
        : grid_(grid), mu_(viscosity) {
        const fftperm_grid fg = to_fftperm_grid(grid);
        fftperm_ctx* raw = nullptr;
        throw_on_error(fftperm_create(&fg, static_cast<double>(viscosity), p_radius, precision, &raw));
        ctx_.reset(raw);
    }

    // ForceSupport(medium, mode) (src/bignonnet/ForceSupport.cpp:9-47) on the GPU.
    std::size_t set_medium(const BinaryMedium& medium, SupportMode mode) {
        if (medium.size() != grid_.total_size()) {
            throw std::invalid_argument("fftperm Context: medium grid mismatch");
        }
        std::uint64_t nb = 0;
        const int m = mode == SupportMode::InterfaceOnly ? FFTPERM_SUPPORT_INTERFACE : FFTPERM_SUPPORT_FULLSOLID;
        throw_on_error(fftperm_set_medium(ctx_.get(), medium.mask().data(), m, &nb));
        nb_ = static_cast<std::size_t>(nb);
        return nb_;
    }

    // ForceSupport::active_voxels (ascending flat indices).
    [[nodiscard]] std::vector<std::size_t> active_voxels() const {
        std::vector<std::uint64_t> a(nb_);
        throw_on_error(fftperm_get_support(ctx_.get(), a.data(), nullptr));
        return std::vector<std::size_t>(a.begin(), a.end());
    }

    [[nodiscard]] fftperm_ctx* get() const noexcept { return ctx_.get(); }
    [[nodiscard]] const Grid3D& grid() const noexcept { return grid_; }
    [[nodiscard]] Scalar viscosity() const noexcept { return mu_; }
    [[nodiscard]] std::size_t num_active_voxels() const noexcept { return nb_; }
    [[nodiscard]] std::size_t size() const noexcept { return 3 * nb_; }  // L = 3 NB

private:
    struct Destroy {
        void operator()(fftperm_ctx* c) const noexcept { fftperm_destroy(c); }
    };
    Grid3D grid_;
    Scalar mu_;
    std::size_t nb_ = 0;
    std::unique_ptr<fftperm_ctx, Destroy> ctx_;
};

// IFFTBackend on the GPU (the reference's SYCL backend sits at the same place,
// src/fft/SYCLFFTBackend.cpp): host fields in, host fields out, one PCIe round trip per call.
// Forward is unnormalised (e^{-i}); inverse is Re(backward C2C) / N, as FFTWBackend
// (src/fft/FFTWBackend.cpp:203-329).  apply_spectral_green_tensor always runs on the device.
class HipFFTBackend final : public IFFTBackend {
public:
    void initialize(const Grid3D& grid, int /*thread_count: host threads do not apply*/) override {
        grid_ = grid;
        const fftperm_grid fg = to_fftperm_grid(grid);
        fftperm_ctx* raw = nullptr;
        // mu and p_radius only shape the context's own Green tensor, which the transform hooks
        // never read: p_radius 0 keeps its set-up to a single term per frequency.
        throw_on_error(fftperm_create(&fg, 1.0, 0, FFTPERM_PRECISION_FP64, &raw));
        ctx_.reset(raw);
        const std::size_t n = grid.total_size();
        real_.assign(3 * n, 0.0);
        spec_a_.assign(6 * n, 0.0);
        spec_b_.assign(6 * n, 0.0);
    }

    void forward(const VectorField3D& real_in, ComplexVectorField3D& complex_out) override {
        require_ready(real_in.size(), complex_out.size());
        const std::size_t n = grid_.total_size();
        const ScalarField3D* comps[3] = {&real_in.x(), &real_in.y(), &real_in.z()};
        for (int c = 0; c < 3; ++c) {
            for (std::size_t i = 0; i < n; ++i) real_[c * n + i] = static_cast<double>((*comps[c])[i]);
        }
        throw_on_error(fftperm_fft_forward(ctx_.get(), real_.data(), spec_a_.data()));
        unpack_spectrum(spec_a_, complex_out);
    }

    void inverse(const ComplexVectorField3D& complex_in, VectorField3D& real_out) override {
        require_ready(real_out.size(), complex_in.size());
        const std::size_t n = grid_.total_size();
        pack_spectrum(complex_in, spec_a_);
        throw_on_error(fftperm_fft_inverse(ctx_.get(), spec_a_.data(), real_.data()));
        ScalarField3D* comps[3] = {&real_out.x(), &real_out.y(), &real_out.z()};
        for (int c = 0; c < 3; ++c) {
            for (std::size_t i = 0; i < n; ++i) (*comps[c])[i] = static_cast<Scalar>(real_[c * n + i]);
        }
    }

    bool apply_spectral_green_tensor(const std::vector<Scalar>& xx, const std::vector<Scalar>& xy,
                                     const std::vector<Scalar>& xz, const std::vector<Scalar>& yy,
                                     const std::vector<Scalar>& yz, const std::vector<Scalar>& zz,
                                     const ComplexVectorField3D& spectral_force,
                                     ComplexVectorField3D& spectral_velocity) override {
        require_ready(spectral_force.size(), spectral_velocity.size());
        const std::size_t n = grid_.total_size();
        const std::vector<Scalar>* t[6] = {&xx, &xy, &xz, &yy, &yz, &zz};
        t6_.resize(6 * n);
        for (int c = 0; c < 6; ++c) {
            if (t[c]->size() != n) throw std::invalid_argument("HipFFTBackend: spectral tensor size mismatch");
            for (std::size_t i = 0; i < n; ++i) t6_[c * n + i] = static_cast<double>((*t[c])[i]);
        }
        pack_spectrum(spectral_force, spec_a_);
        throw_on_error(fftperm_spectral_green(ctx_.get(), t6_.data(), spec_a_.data(), spec_b_.data()));
        unpack_spectrum(spec_b_, spectral_velocity);
        return true;
    }

    [[nodiscard]] double normalization_factor() const noexcept override {
        return 1.0 / static_cast<double>(grid_.total_size());
    }
    [[nodiscard]] const Grid3D& grid() const noexcept override { return grid_; }
    [[nodiscard]] std::string device_info() const override {
        char buf[256] = {0};
        if (fftperm_device_info(buf, sizeof(buf)) != FFTPERM_OK) return {};
        return std::string("libfftperm: ") + buf;
    }

private:
    struct Destroy {
        void operator()(fftperm_ctx* c) const noexcept { fftperm_destroy(c); }
    };

    void require_ready(std::size_t a, std::size_t b) const {
        if (!ctx_) throw std::runtime_error("HipFFTBackend used before initialize()");
        if (a != grid_.total_size() || b != grid_.total_size()) {
            throw std::invalid_argument("HipFFTBackend: field size does not match the initialized grid");
        }
    }

    void pack_spectrum(const ComplexVectorField3D& f, std::vector<double>& out) const {
        const std::size_t n = grid_.total_size();
        const Array3D<ScalarComplex>* comps[3] = {&f.x(), &f.y(), &f.z()};
        for (int c = 0; c < 3; ++c) {
            for (std::size_t i = 0; i < n; ++i) {
                const ScalarComplex v = (*comps[c])[i];
                out[2 * (c * n + i)] = static_cast<double>(v.real());
                out[2 * (c * n + i) + 1] = static_cast<double>(v.imag());
            }
        }
    }

    void unpack_spectrum(const std::vector<double>& in, ComplexVectorField3D& f) const {
        const std::size_t n = grid_.total_size();
        Array3D<ScalarComplex>* comps[3] = {&f.x(), &f.y(), &f.z()};
        for (int c = 0; c < 3; ++c) {
            for (std::size_t i = 0; i < n; ++i) {
                (*comps[c])[i] = ScalarComplex(static_cast<Scalar>(in[2 * (c * n + i)]),
                                               static_cast<Scalar>(in[2 * (c * n + i) + 1]));
            }
        }
    }

    Grid3D grid_{1, 1, 1};
    std::unique_ptr<fftperm_ctx, Destroy> ctx_;
    std::vector<double> real_, spec_a_, spec_b_, t6_;
};

// The ComputeBackend::GPU case of create_fft_backend (src/fft/FFTBackendFactory.cpp:13-34):
//     case ComputeBackend::GPU: return permeability::hip::make_gpu_fft_backend();
inline std::unique_ptr<IFFTBackend> make_gpu_fft_backend() { return std::make_unique<HipFFTBackend>(); }

// VariationalOperator (src/bignonnet/VariationalOperator.cpp:19-41) on the GPU: the support map,
// Green tensor and transforms stay on the device; each apply moves the compact vector (L = 3 NB)
// in and out.  Usable with the reference's own MinresSolver.
class HipVariationalOperator final : public ILinearOperator {
public:
    explicit HipVariationalOperator(Context& ctx) : ctx_(&ctx) {}

    [[nodiscard]] std::size_t size() const override { return ctx_->size(); }

    void apply(const std::vector<Scalar>& in, std::vector<Scalar>& out) const override {
        if (in.size() != size()) {
            throw std::invalid_argument("VariationalOperator::apply input size mismatch");
        }
        DoubleIn x(in);
        DoubleOut y(out, size());
        throw_on_error(fftperm_operator_apply(ctx_->get(), x.data(), y.data()));
        y.commit();
    }

    [[nodiscard]] Context& context() const noexcept { return *ctx_; }

private:
    Context* ctx_;
};

// MinresSolver (src/solver/MinresSolver.cpp:46-218) with the whole iteration on the GPU: same
// constructor checks, same solve() arguments, same SolverResult.  The device loop does not
// return to the host between iterations, so a progress callback is invoked after the solve, for
// every progress_interval-th entry of the residual history.
class HipMinresSolver {
public:
    using ProgressCallback = std::function<void(std::size_t, Scalar)>;

    HipMinresSolver(Scalar tolerance, std::size_t max_iterations)
        : tolerance_(tolerance), max_iterations_(max_iterations) {
        if (tolerance_ <= Scalar(0)) {
            throw std::invalid_argument("MinresSolver tolerance must be positive");
        }
    }

    [[nodiscard]] SolverResult solve(const HipVariationalOperator& op, const std::vector<Scalar>& rhs,
                                     const std::vector<Scalar>& initial_guess = {}, int progress_interval = 0,
                                     ProgressCallback progress_callback = {}) const {
        const std::size_t n = op.size();
        if (rhs.size() != n) throw std::invalid_argument("MinresSolver::solve rhs size mismatch");
        if (!initial_guess.empty() && initial_guess.size() != n) {
            throw std::invalid_argument("MinresSolver::solve initial guess size mismatch");
        }
        DoubleIn b(rhs);
        std::optional<DoubleIn> x0;
        if (!initial_guess.empty()) x0.emplace(initial_guess);
        SolverResult result;
        DoubleOut x(result.solution, n);
        std::vector<double> history(max_iterations_ + 1, 0.0);
        fftperm_solve_info info{};
        throw_on_error(fftperm_minres(op.context().get(), b.data(), x0 ? x0->data() : nullptr,
                                      static_cast<double>(tolerance_), max_iterations_, x.data(), history.data(),
                                      &info));
        x.commit();
        result.converged = info.converged != 0;
        result.iterations = static_cast<std::size_t>(info.iterations);
        result.final_rel_residual = static_cast<Scalar>(info.final_rel_residual);
        result.residual_history.reserve(info.history_len);
        for (std::uint64_t i = 0; i < info.history_len; ++i) {
            result.residual_history.push_back(static_cast<Scalar>(history[i]));
        }
        if (progress_interval > 0 && progress_callback) {
            for (std::size_t it = static_cast<std::size_t>(progress_interval); it < result.residual_history.size();
                 it += static_cast<std::size_t>(progress_interval)) {
                progress_callback(it, result.residual_history[it]);
            }
        }
        return result;
    }

private:
    Scalar tolerance_;
    std::size_t max_iterations_;
};

// PermeabilityRunner's DirectionSolveResult (src/app/PermeabilityRunner.cpp:173-242).
struct DirectionSolve {
    std::size_t col = 0;
    SolverResult solve_result;
    Real3 avg_velocity{Scalar(0), Scalar(0), Scalar(0)};
    std::optional<VectorField3D> velocity_field;
};

// solve_one_direction: VariationalRHS::build, MinresSolver::solve (x0 empty unless given),
// VelocityRecovery::recover and average_over_domain, all on the device of ctx.
inline DirectionSolve solve_one_direction(Context& ctx, std::size_t col, const Real3& grad, Scalar tolerance,
                                          std::size_t max_iterations, bool want_velocity_field,
                                          const std::vector<Scalar>& initial_guess = {}) {
    if (ctx.num_active_voxels() == 0) {
        throw std::runtime_error("Support set is empty; cannot run variational solve");
    }
    const double g[3] = {static_cast<double>(grad[0]), static_cast<double>(grad[1]), static_cast<double>(grad[2])};
    std::vector<Scalar> b;
    {
        DoubleOut bo(b, ctx.size());
        throw_on_error(fftperm_rhs(ctx.get(), g, bo.data()));
        bo.commit();
    }
    HipVariationalOperator op(ctx);
    DirectionSolve out;
    out.col = col;
    out.solve_result = HipMinresSolver(tolerance, max_iterations).solve(op, b, initial_guess);
    DoubleIn x(out.solve_result.solution);
    const std::size_t n = ctx.grid().total_size();
    std::vector<double> u(want_velocity_field ? 3 * n : 0);
    double avg[3] = {0.0, 0.0, 0.0};
    throw_on_error(fftperm_recover(ctx.get(), g, x.data(), u.empty() ? nullptr : u.data(), avg));
    out.avg_velocity = Real3{static_cast<Scalar>(avg[0]), static_cast<Scalar>(avg[1]), static_cast<Scalar>(avg[2])};
    if (want_velocity_field) {
        VectorField3D v(ctx.grid());
        for (std::size_t i = 0; i < n; ++i) {
            v.x()[i] = static_cast<Scalar>(u[i]);
            v.y()[i] = static_cast<Scalar>(u[n + i]);
            v.z()[i] = static_cast<Scalar>(u[2 * n + i]);
        }
        out.velocity_field = std::move(v);
    }
    return out;
}

// gradient=all: one unit gradient per column, colK = -mu <v>, PermeabilityTensor::set_column
// negates it again, so K[r][c] = +mu <v_r>^(c) (src/app/PermeabilityRunner.cpp:757-777).
inline PermeabilityTensor solve_permeability_tensor(Context& ctx, Scalar tolerance, std::size_t max_iterations,
                                                    std::array<DirectionSolve, 3>* per_direction = nullptr) {
    PermeabilityTensor K;
    for (std::size_t col = 0; col < 3; ++col) {
        Real3 grad{Scalar(0), Scalar(0), Scalar(0)};
        grad[col] = Scalar(1);
        DirectionSolve d = solve_one_direction(ctx, col, grad, tolerance, max_iterations, false);
        const Scalar mu = ctx.viscosity();
        const Real3 colK{-mu * d.avg_velocity[0], -mu * d.avg_velocity[1], -mu * d.avg_velocity[2]};
        K.set_column(col, colK);
        if (per_direction) (*per_direction)[col] = std::move(d);
    }
    return K;
}

}  // namespace permeability::hip
