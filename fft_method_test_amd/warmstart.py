"""Learned warm start for the Bignonnet MINRES solves (no counterpart in the reference).

A 4-level 3D U-Net (16 base channels, BF16 weights/activations, FP32 accumulation; HIP kernels in
csrc/unet*.hip) maps a periodic fluid mask to the interface force density of the x-direction
variational problem.  It is trained on solver-generated (medium, force density) pairs by its own
forward / backward / Adam kernels, and its prediction seeds MINRES; MINRES stays the judge of
the error (fftperm_solve_tensor_warm drops a guess that does not lower the FP64 true residual).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib, geometry
from ._lib import check

# the flat parameter layout of the network, per layer (taps, cin0, cin1, cout)
LAYERS = ((27, 8, 0, 16), (27, 16, 0, 16), (27, 16, 0, 32), (27, 32, 0, 32), (27, 32, 0, 64), (27, 64, 0, 64),
          (27, 64, 0, 128), (27, 128, 0, 128), (27, 128, 64, 64), (27, 64, 0, 64), (27, 64, 32, 32),
          (27, 32, 0, 32), (27, 32, 16, 16), (27, 16, 0, 16), (1, 16, 0, 16))


def param_count():
    return sum(co * t * (c0 + c1) + co for t, c0, c1, co in LAYERS)


class WarmStartNet:
    """Handle of the device network (fftperm_unet_*)."""

    def __init__(self, seed=0):
        L = _lib.load()
        h = C.c_void_p()
        check(L.fftperm_unet_create(int(seed), C.byref(h)))
        self._h = h
        n = C.c_uint64()
        check(L.fftperm_unet_num_params(self._h, C.byref(n)))
        self.num_params = int(n.value)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().fftperm_unet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        check(_lib.load().fftperm_unet_get_params(self._h, p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError("parameter vector size mismatch")
        check(_lib.load().fftperm_unet_set_params(self._h, p))

    def save(self, path, bf16=False):
        """FP32 parameters, or (bf16=True) their BF16 bit patterns -- the precision the device
        network computes with -- at half the size."""
        p = self.get_params()
        if bf16:
            bits = (p.view(np.uint32) + 0x7FFF + ((p.view(np.uint32) >> 16) & 1)) >> 16
            np.savez(path, params_bf16=bits.astype(np.uint16), layers=np.asarray(LAYERS, dtype=np.int32))
        else:
            np.savez(path, params=p, layers=np.asarray(LAYERS, dtype=np.int32))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            if not np.array_equal(z["layers"], np.asarray(LAYERS, dtype=np.int32)):
                raise ValueError("saved network has a different architecture")
            if "params_bf16" in z:
                p = (z["params_bf16"].astype(np.uint32) << 16).view(np.float32)
            else:
                p = z["params"]
            self.set_params(p)

    def loss_grad(self, mask_dev, target_dev, n, lr=0.0, want_grad=False):
        """mask_dev: torch uint8 (n^3, fluid = 1) on the GPU; target_dev: torch float32 (n^3, 3).
        Returns (loss, grads or None); lr > 0 also takes one Adam step."""
        import torch

        torch.cuda.synchronize()
        loss = C.c_double()
        g = np.zeros(self.num_params, dtype=np.float32) if want_grad else None
        check(_lib.load().fftperm_unet_loss_grad(self._h, C.c_void_p(mask_dev.data_ptr()),
                                                 C.c_void_p(target_dev.data_ptr()), int(n), float(lr),
                                                 C.byref(loss), g.ctypes.data_as(C.c_void_p) if want_grad else None))
        return float(loss.value), g

    def predict(self, mask_dev, n):
        import torch

        torch.cuda.synchronize()
        out = torch.empty((n ** 3, 3), dtype=torch.float32, device=mask_dev.device)
        check(_lib.load().fftperm_unet_predict(self._h, C.c_void_p(mask_dev.data_ptr()), int(n),
                                               C.c_void_p(out.data_ptr())))
        return out


def solve_tensor_warm(green, net: WarmStartNet, tolerance, max_iterations):
    """fftperm_solve_tensor seeded by the network: (K 3x3, infos, used_guess[3])."""
    K = np.zeros(9)
    infos = (_lib.SolveInfo * 3)()
    used = (C.c_int * 3)()
    check(_lib.load().fftperm_solve_tensor_warm(green._h, net._h, float(tolerance), int(max_iterations), K,
                                                C.cast(infos, C.c_void_p), C.cast(used, C.c_void_p)))
    return K.reshape(3, 3), list(infos), [int(u) for u in used]


def random_medium(n, rng):
    """Training / evaluation media at the voxel scale of the benchmark problems: overlapping
    radius-8 spheres or prolate ellipsoids (semi-axes 12, 4, 4, random orientation), half each, at
    a porosity drawn from U[0.3, 0.8]."""
    phi = float(rng.uniform(0.3, 0.8))
    seed = int(rng.integers(0, 2 ** 31 - 1))
    if rng.uniform() < 0.5:
        return geometry.boolean_spheres(n, 8.0, phi, seed)
    return geometry.boolean_ellipsoids(n, (12.0, 4.0, 4.0), phi, seed)


def force_density_x(green, support, tol, maxit):
    """Direction-x force density of the variational problem on the grid (n^3, 3) float32, from
    the GPU solver (RHS + MINRES, FP64)."""
    from . import permeability as P

    b = P.VariationalRHS(support, green, None).build([1.0, 0.0, 0.0])
    res = P.MinresSolver(tol, maxit).solve(P.VariationalOperator(support, green), b)
    N = green.grid().total_size()
    field = np.zeros((N, 3), dtype=np.float32)
    field[support.active_voxels().astype(np.int64)] = res.solution.reshape(-1, 3)
    return field, res


def generate_pairs(n, count, seed, tol=1e-6, maxit=20000, log=None):
    """`count` (fluid mask, force density normalised to unit RMS on the support) pairs at n^3
    from the FP64 GPU solver."""
    from . import permeability as P

    rng = np.random.default_rng(seed)
    g = P.Grid3D(n, n, n)
    green = P.GreenOperator(g)
    pairs = []
    t0 = time.perf_counter()
    for i in range(count):
        mask = random_medium(n, rng).reshape(-1)
        sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
        if sup.num_active_voxels() == 0:
            continue
        field, res = force_density_x(green, sup, tol, maxit)
        # unit RMS over the support: the network learns the shape, the solve fixes the scale
        # (minimal-residual step along the prediction)
        rms = float(np.sqrt(np.mean(field[sup.active_voxels().astype(np.int64)] ** 2)))
        if rms > 0:
            field /= rms
        pairs.append((mask.astype(np.uint8), field))
        if log and (i + 1) % max(1, count // 10) == 0:
            log(f"[warmstart] pairs {i + 1}/{count} at {n}^3 ({time.perf_counter() - t0:.1f} s, last "
                f"{res.iterations} iterations)")
    green.close()
    return pairs


def train(net: WarmStartNet, pairs, n, epochs, lr=1e-3, seed=0, log=None):
    """Adam over the pairs (one sample per step, shuffled each epoch); returns the loss history."""
    import torch

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    masks = [torch.from_numpy(m).to(dev) for m, _ in pairs]
    targets = [torch.from_numpy(np.ascontiguousarray(f)).to(dev) for _, f in pairs]
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        tot = 0.0
        for i in order:
            l, _ = net.loss_grad(masks[i], targets[i], n, lr=lr)
            tot += l
        hist.append(tot / max(1, len(pairs)))
        if log:
            log(f"[warmstart] epoch {ep + 1}/{epochs} mean loss {hist[-1]:.4e}")
    return hist


def train_residual(net: WarmStartNet, masks, n, epochs, lr=2e-4, seed=0, log=None):
    """Train on the quantity MINRES sees: the relative residual of the scaled guess,
    loss = min_alpha ||b - alpha A p||^2 / ||b||^2 (p = the prediction gathered on the support,
    b the direction-x RHS, A the FP64 variational operator).  At the optimal alpha its gradient
    with respect to p is -2 alpha A r / ||b||^2, r = b - alpha A p (A is symmetric).  The network's
    squared-error kernel delivers a prescribed output gradient g when trained against the target
    pred - (3 count / 2) g, count = support voxels, which is how the step is taken.  Returns the
    per-epoch mean relative residual (the square root of the loss)."""
    import torch

    from . import permeability as P

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    g = P.Grid3D(n, n, n)
    green = P.GreenOperator(g)
    samples = []
    for m in masks:
        sup = P.ForceSupport(P.BinaryMedium(g, m), P.SUPPORT_INTERFACE, green)
        if sup.num_active_voxels() == 0:
            continue
        b = P.VariationalRHS(sup, green, None).build([1.0, 0.0, 0.0])
        samples.append((np.ascontiguousarray(m, dtype=np.uint8), sup.active_voxels().astype(np.int64), b))
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(len(samples)):
            m, act, b = samples[i]
            nb_ = np.zeros(1, dtype=np.uint64)
            check(_lib.load().fftperm_set_medium(green._h, m, _lib.SUPPORT_INTERFACE,
                                                 nb_.ctypes.data_as(C.POINTER(C.c_uint64))))
            mdev = torch.from_numpy(m).to(dev)
            pred = net.predict(mdev, n).cpu().numpy()
            p = pred[act].astype(np.float64).reshape(-1)
            Ap = _apply(green, p)
            bb = float(b @ b)
            apap = float(Ap @ Ap)
            alpha = float(b @ Ap) / apap if apap > 0 else 0.0
            r = b - alpha * Ap
            tot += np.sqrt(float(r @ r) / bb)
            grad_p = (-2.0 * alpha / bb) * _apply(green, r)
            gfield = np.zeros_like(pred)
            gfield[act] = grad_p.reshape(-1, 3)
            target = (pred - 1.5 * act.size * gfield).astype(np.float32)
            net.loss_grad(mdev, torch.from_numpy(target).to(dev), n, lr=lr)
        hist.append(tot / max(1, len(samples)))
        if log:
            log(f"[warmstart] residual epoch {ep + 1}/{epochs} mean relative residual of the guess {hist[-1]:.4f}")
    green.close()
    return hist


def _apply(green, x):
    y = np.zeros_like(x)
    check(_lib.load().fftperm_operator_apply(green._h, np.ascontiguousarray(x, dtype=np.float64), y))
    return y
