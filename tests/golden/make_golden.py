"""Write the golden fixtures of tests/golden/ from the reference's own compiled code.

The vectors come from oracle/_ref/libref.so (built by oracle/Makefile from the
reference sources; FFTW is absent, so its FFT backend runs its identity fallback
and GreenOperator::apply is a pointwise multiplication by the reference's Green
tensor).  Inputs are seeded; outputs are what the reference computes for them.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def random_mask(shape, porosity, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(int(np.prod(shape))) < porosity).astype(np.uint8)


def bignonnet_case(name, shape, porosity, seed, support, mu, p, h):
    mask = random_mask(shape, porosity, seed)
    r = oracle.Ref(shape, mask, support=support, h=h, mu=mu, p_radius=p)
    act, v2l = r.support()
    rng = np.random.default_rng(seed + 100)
    grad = np.array([1.0, -0.5, 0.25])
    x = rng.standard_normal(3 * r.nb)
    f = rng.standard_normal(3 * r.N)
    b = r.rhs(grad)
    sol, info, hist = r.minres(b, 1e-10, 300)
    avg, avg_b, u = r.recover(grad, sol, want_field=True)
    K = np.zeros(9)
    oracle.ref_set_column_k(mu, 0, avg, K)
    return {
        f"{name}_shape": np.array(shape),
        f"{name}_params": np.array([porosity, seed, support, mu, p, h]),
        f"{name}_mask": mask,
        f"{name}_active": act,
        f"{name}_v2l": v2l,
        f"{name}_green": r.green_tensor(),
        f"{name}_f0": r.background_force(grad),
        f"{name}_f": f,
        f"{name}_green_apply": r.green_apply(f),
        f"{name}_rhs": b,
        f"{name}_x": x,
        f"{name}_Ax": r.apply(x),
        f"{name}_minres_x": sol,
        f"{name}_minres_hist": hist,
        f"{name}_minres_iters": np.array([info["iterations"], info["converged"]]),
        f"{name}_recover_u": u,
        f"{name}_recover_avg": avg,
        f"{name}_Kcol0": K,
        f"{name}_residual_norm": np.array([r.residual_norm(sol, b)]),
        f"{name}_symmetry_gap": np.array([r.symmetry_gap(x, f[: x.size] if f.size >= x.size else x)]),
    }


def brinkman_case(name, shape, porosity, seed, penalty, maxit):
    mask = random_mask(shape, porosity, seed)
    ids = np.where(mask == 1, 1, 2).astype(np.uint32)
    U, info, u = oracle.ref_brinkman_solve(shape, ids, 1, 1.0, penalty, [1.0, 0.0, 0.0], relaxation=0.9,
                                           tol=1e-8, maxit=maxit, want_field=True)
    K, its, res = oracle.ref_brinkman_tensor(shape, ids, 1, 1.0, penalty, relaxation=0.9, tol=1e-8, maxit=maxit)
    return {
        f"{name}_shape": np.array(shape),
        f"{name}_params": np.array([porosity, seed, penalty, maxit]),
        f"{name}_mask": mask,
        f"{name}_U": U,
        f"{name}_u": u,
        f"{name}_info": np.array([info["iterations"], info["residual"], info["converged"]]),
        f"{name}_K": K,
        f"{name}_K_iters": np.array(its),
        f"{name}_K_res": np.array(res),
    }


def main():
    oracle.build(ref=True)
    data = {}
    data.update(bignonnet_case("b1", (6, 5, 4), 0.7, 1, oracle.SUPPORT_INTERFACE, 1.0, 2, 1.0))
    data.update(bignonnet_case("b2", (4, 4, 4), 0.6, 2, oracle.SUPPORT_FULLSOLID, 1.3, 1, 0.5))
    data.update(bignonnet_case("b3", (8, 7, 6), 0.75, 3, oracle.SUPPORT_INTERFACE, 0.8, 2, 2.0))
    data.update(brinkman_case("k1", (6, 5, 4), 0.7, 4, 1e4, 40))
    data.update(brinkman_case("k2", (8, 8, 8), 0.6, 5, 1e2, 25))
    out = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(out, **data)
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
