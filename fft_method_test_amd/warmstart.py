"""Learned warm start for the Bignonnet MINRES solves (no counterpart in the reference).

A 4-level 3D U-Net (16 base channels, BF16 weights/activations, FP32 accumulation; HIP kernels in
csrc/unet*.hip) maps a periodic fluid mask to the interface force densities of the three unit load
directions of the variational problem in one forward pass (output channels 3d..3d+2 for direction d).  It is trained on solver-generated (medium, force density) pairs by its own
forward / backward / Adam kernels, and its prediction seeds MINRES; MINRES stays the judge of
the error (fftperm_solve_tensor_warm drops a guess that does not lower the FP64 true residual).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib, geometry
from ._lib import check

# predicted channels: three load directions x three components
OUTPUTS = 9
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
            np.savez(path, params_bf16=bits.astype(np.uint16), layers=np.asarray(LAYERS, dtype=np.int32),
                     outputs=np.int32(OUTPUTS))
        else:
            np.savez(path, params=p, layers=np.asarray(LAYERS, dtype=np.int32), outputs=np.int32(OUTPUTS))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            if not np.array_equal(z["layers"], np.asarray(LAYERS, dtype=np.int32)):
                raise ValueError("saved network has a different architecture")
            if "outputs" not in z or int(z["outputs"]) != OUTPUTS:
                raise ValueError("saved network predicts a different set of outputs")
            if "params_bf16" in z:
                p = (z["params_bf16"].astype(np.uint32) << 16).view(np.float32)
            else:
                p = z["params"]
            self.set_params(p)

    def loss_grad(self, mask_dev, target_dev, n, lr=0.0, want_grad=False):
        """mask_dev: torch uint8 (n^3, fluid = 1) on the GPU; target_dev: torch float32 (n^3, 9): the
        force densities of load directions x, y, z (three components each).
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
        out = torch.empty((n ** 3, OUTPUTS), dtype=torch.float32, device=mask_dev.device)
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


def medium_for_seed(n, seed):
    """The medium of one training / held-out seed (flat uint8 n^3, fluid = 1)."""
    return np.ascontiguousarray(random_medium(n, np.random.default_rng(seed)).reshape(-1), dtype=np.uint8)


def _media_worker(args):
    n, seed = args
    return seed, medium_for_seed(n, seed)


def media(n, seeds, workers=1):
    """Yields (seed, mask) in seed order; workers > 1 generates them in spawned processes (the
    ellipsoid packs take ~1 s each at 128^3 on one core)."""
    if workers <= 1:
        for sd in seeds:
            yield sd, medium_for_seed(n, sd)
        return
    import multiprocessing as mp

    with mp.get_context("spawn").Pool(workers) as pool:
        for out in pool.imap(_media_worker, [(n, sd) for sd in seeds], chunksize=2):
            yield out


def density_targets(green, support, tol, maxit):
    """The three load directions' force densities on the support (NB x 9 float32: direction d's
    components in columns 3d..3d+2), each scaled to unit RMS over the support (the network learns
    the shape; the solve fixes the scale by the minimal-residual step along it), from the GPU
    solver (RHS + MinresSolver of the context's precision).  Returns (densities, iterations[3])."""
    from . import permeability as P

    op = P.VariationalOperator(support, green)
    out = np.zeros((support.num_active_voxels(), OUTPUTS), dtype=np.float32)
    its = []
    for d in range(3):
        g = [0.0, 0.0, 0.0]
        g[d] = 1.0
        b = P.VariationalRHS(support, green, None).build(g)
        res = P.MinresSolver(tol, maxit).solve(op, b)
        x = res.solution.reshape(-1, 3)
        rms = float(np.sqrt(np.mean(x * x)))
        out[:, 3 * d:3 * d + 3] = x / rms if rms > 0 else x
        its.append(int(res.iterations))
    return out, its


class TrainingSet:
    """Solver-generated (medium, densities) pairs held on the GPU: per pair the mask (n^3 uint8),
    the support voxels (int64) and the densities on them (NB x 9 float32)."""

    def __init__(self, n):
        self.n = n
        self.items = []

    def __len__(self):
        return len(self.items)

    def add(self, mask, active, dens):
        import torch

        dev = torch.device("cuda:0")
        self.items.append((torch.from_numpy(mask).to(dev), torch.from_numpy(active.astype(np.int64)).to(dev),
                           torch.from_numpy(dens).to(dev)))

    def target(self, i, buf):
        """Writes pair i's target field ([n^3][9], zero off the support) into buf."""
        _, act, dens = self.items[i]
        buf.zero_()
        buf[act] = dens
        return buf


def generate(n, seeds, tol=1e-5, maxit=20000, workers=1, precision=None, log=None, into=None):
    """Training pairs for the given seeds at n^3: media (spawned workers), then the GPU solves."""
    from . import permeability as P

    prec = _lib.PRECISION_FP64 if precision is None else precision
    data = TrainingSet(n) if into is None else into
    g = P.Grid3D(n, n, n)
    green = P.GreenOperator(g, 1.0, 2, prec)
    t0 = time.perf_counter()
    seeds = list(seeds)
    for k, (sd, mask) in enumerate(media(n, seeds, workers)):
        sup = P.ForceSupport(P.BinaryMedium(g, mask), P.SUPPORT_INTERFACE, green)
        if sup.num_active_voxels() == 0:
            continue
        dens, its = density_targets(green, sup, tol, maxit)
        data.add(mask, sup.active_voxels().astype(np.int64), dens)
        if log and ((k + 1) % 25 == 0 or k + 1 == len(seeds)):
            log(f"[warmstart] pairs {k + 1}/{len(seeds)} at {n}^3 ({time.perf_counter() - t0:.1f} s, last seed {sd} "
                f"iterations {its})")
    green.close()
    return data


def train(net: WarmStartNet, data: TrainingSet, epochs, lr=1e-3, lr_min=2e-5, seed=0, log=None, deadline=None):
    """Adam over the pairs, one pair per step, shuffled each epoch; the learning rate follows a
    cosine from lr to lr_min over all steps.  Stops early (after finishing a step) past `deadline`
    (time.perf_counter() value).  Returns the per-epoch mean loss."""
    import math

    import torch

    rng = np.random.default_rng(seed)
    n = data.n
    buf = torch.zeros((n ** 3, OUTPUTS), dtype=torch.float32, device="cuda:0")
    total = max(1, epochs * len(data))
    step = 0
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        tot, cnt = 0.0, 0
        for i in rng.permutation(len(data)):
            a = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * step / total))
            l, _ = net.loss_grad(data.items[i][0], data.target(int(i), buf), n, lr=a)
            tot += l
            cnt += 1
            step += 1
            if deadline is not None and time.perf_counter() > deadline:
                break
        hist.append(tot / max(1, cnt))
        if log:
            log(f"[warmstart] epoch {ep + 1}/{epochs} mean loss {hist[-1]:.4e} ({time.perf_counter() - t0:.0f} s)")
        if deadline is not None and time.perf_counter() > deadline:
            if log:
                log("[warmstart] time budget reached")
            break
    return hist
