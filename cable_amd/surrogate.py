"""Host side of the canopy surrogate (include/cable_mlp.h): parameter
initialisation, normalisation statistics and the training loop that drives
the GPU kernels of cable_amd/csrc/mlp.hip.  Every forward, backward and Adam
step runs on the GPU; the host only draws the sample order and reads losses.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .gpu import CableError, lib as _gpu_lib

NIN, HID, NOUT = 24, 128, 8
SHAPES = [(NIN, HID), (HID, HID), (HID, HID), (HID, NOUT)]
NPARAM = sum(i * o + o for i, o in SHAPES)

_sig_done = False


def lib():
    global _sig_done
    L = _gpu_lib()
    if not _sig_done:
        vp, i64, c_int, c_float = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float
        sig = {
            "cable_mlp_nparam": ([], c_int),
            "cable_mlp_last_error": ([], ctypes.c_char_p),
            "cable_mlp_create": ([ctypes.POINTER(vp), c_int], c_int),
            "cable_mlp_destroy": ([vp], c_int),
            "cable_mlp_set_params": ([vp, vp], c_int),
            "cable_mlp_get_params": ([vp, vp], c_int),
            "cable_mlp_set_norm": ([vp, vp], c_int),
            "cable_mlp_forward": ([vp, vp, c_int, vp], c_int),
            "cable_mlp_gradient": ([vp, vp, vp, c_int, vp], c_int),
            "cable_mlp_get_grads": ([vp, vp], c_int),
            "cable_mlp_load_samples": ([vp, vp, vp, i64], c_int),
            "cable_mlp_use_samples": ([vp, vp, vp, i64], c_int),
            "cable_mlp_train": ([vp, vp, i64, c_float, vp], c_int),
            "cable_mlp_train_shuffled": ([vp, ctypes.c_uint64, i64, i64, c_float, vp], c_int),
            "cable_mlp_shuffled_order": ([ctypes.c_uint64, i64, i64, i64, vp], c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        if L.cable_mlp_nparam() != NPARAM:
            raise CableError("libcable_gpu.so has a different surrogate layout")
        _sig_done = True
    return L


def _check(rc):
    if rc != 0:
        raise CableError(lib().cable_mlp_last_error().decode())


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise CableError(f"expected shape {shape}, got {a.shape}")
    return a


def init_params(seed=0):
    """He-uniform weights (hard-swish layers), zero biases, a half-scale output layer."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (i, o) in enumerate(SHAPES):
        lim = np.sqrt(6.0 / i) * (0.5 if l == 3 else 1.0)
        parts += [rng.uniform(-lim, lim, i * o), np.zeros(o)]
    return np.concatenate(parts).astype(np.float32)


def norm_stats(x, y):
    """mu_x, sd_x, mu_y, sd_y of raw samples (sd floored so constant
    columns stay finite), packed as the library expects."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    mx, sx = x.mean(0), x.std(0)
    my, sy = y.mean(0), y.std(0)
    sx = np.where(sx > 1e-6 * np.maximum(1.0, np.abs(mx)), sx, 1.0)
    sy = np.where(sy > 1e-6 * np.maximum(1.0, np.abs(my)), sy, 1.0)
    return np.concatenate([mx, sx, my, sy]).astype(np.float32)


class MLP:
    """One network on the GPU with buffers for batches of ``batch`` rows."""

    def __init__(self, batch=2048, params=None, norm=None):
        self.batch = int(batch)
        h = ctypes.c_void_p()
        _check(lib().cable_mlp_create(ctypes.byref(h), self.batch))
        self.h = h
        self.set_params(init_params() if params is None else params)
        if norm is not None:
            self.set_norm(norm)

    def close(self):
        if getattr(self, "h", None):
            lib().cable_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = _f32(p, (NPARAM,))
        _check(lib().cable_mlp_set_params(self.h, p.ctypes.data))

    def params(self):
        out = np.empty(NPARAM, dtype=np.float32)
        _check(lib().cable_mlp_get_params(self.h, out.ctypes.data))
        return out

    def set_norm(self, norm):
        self.norm = _f32(norm, (2 * NIN + 2 * NOUT,))
        _check(lib().cable_mlp_set_norm(self.h, self.norm.ctypes.data))

    def forward(self, x):
        """Network output for normalised inputs x (n x 24), n <= batch."""
        x = _f32(x)
        y = np.empty((x.shape[0], NOUT), dtype=np.float32)
        _check(lib().cable_mlp_forward(self.h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def gradient(self, x, t):
        x, t = _f32(x), _f32(t)
        loss = np.zeros(1, dtype=np.float32)
        _check(lib().cable_mlp_gradient(self.h, x.ctypes.data, t.ctypes.data, x.shape[0], loss.ctypes.data))
        g = np.empty(NPARAM, dtype=np.float32)
        _check(lib().cable_mlp_get_grads(self.h, g.ctypes.data))
        return float(loss[0]), g

    def load_samples(self, x, y):
        self._xs, self._ys = _f32(x), _f32(y)
        if self._xs.shape[1] != NIN or self._ys.shape[1] != NOUT or self._xs.shape[0] != self._ys.shape[0]:
            raise CableError("samples must be (n, 24) and (n, 8)")
        _check(lib().cable_mlp_load_samples(self.h, self._xs.ctypes.data, self._ys.ctypes.data,
                                            self._xs.shape[0]))
        self.nsamples = self._xs.shape[0]

    def use_device_samples(self, x_ptr, y_ptr, n):
        _check(lib().cable_mlp_use_samples(self.h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), int(n)))
        self.nsamples = int(n)

    def train(self, epochs=1.0, lr=1e-3, seed=0, lr_final=None):
        """Adam over shuffled batches of the loaded samples for ``epochs``
        passes; the learning rate decays linearly to lr_final (default lr/10)
        in chunks.  The device draws the order itself (each epoch a
        permutation keyed by seed and epoch; see shuffled_order).  Returns
        the loss of every batch."""
        n, B = self.nsamples, self.batch
        nb = max(1, int(round(epochs * n / B)))
        lr_final = lr / 10.0 if lr_final is None else lr_final
        losses = np.zeros(nb, dtype=np.float32)
        done, chunk = 0, 512
        while done < nb:
            k = min(chunk, nb - done)
            rate = lr + (lr_final - lr) * (done + 0.5 * k) / nb
            _check(lib().cable_mlp_train_shuffled(self.h, int(seed), done, k, float(rate),
                                                  losses[done:].ctypes.data))
            done += k
        return losses

    def train_order(self, order, lr):
        """Adam steps over batches in a given order of sample indices
        (len(order) a multiple of the batch).  Returns the batch losses."""
        order = np.ascontiguousarray(order, dtype=np.int64)
        k = order.size // self.batch
        losses = np.zeros(k, dtype=np.float32)
        _check(lib().cable_mlp_train(self.h, order.ctypes.data, k, float(lr), losses.ctypes.data))
        return losses


def shuffled_order(seed, n, first, count):
    """Sample indices at positions first .. first+count-1 of the training
    order over n samples (epoch e = position // n is a permutation of
    range(n) keyed by seed and e)."""
    out = np.empty(int(count), dtype=np.int64)
    _check(lib().cable_mlp_shuffled_order(int(seed), int(n), int(first), int(count), out.ctypes.data))
    return out


def capture_samples(n_tiles=262144, nsteps=64, seed=None, order=True):
    """Run the plain mixed-precision step on a fresh synthetic workload and
    keep one training sample per vegetated tile-step on the GPU.  Returns
    the context (which owns the samples) and the sample count."""
    from . import gpu
    from .workload import DEFAULT_SEED, Workload
    wl = Workload(n_tiles, nsteps, seed=(DEFAULT_SEED + 1) if seed is None else seed)
    o = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, tair=wl.t0) if order else None
    g = gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=o)
    n = g.run_capture(0, nsteps, wl.dels, n_tiles * nsteps)
    return g, n


def sample_stats(g, n, chunks=32, rows=16384):
    """Normalisation statistics from evenly spaced blocks of the captured
    samples (which are stored in step order)."""
    xs, ys = [], []
    rows = min(rows, max(1, n // chunks))
    for c in range(chunks):
        first = min(n - rows, (c * n) // chunks)
        x, y = g.copy_samples(first, rows)
        xs.append(x)
        ys.append(y)
    return norm_stats(np.concatenate(xs), np.concatenate(ys))


def train_on_gpu(n_tiles=262144, nsteps=64, epochs=2.0, batch=2048, lr=2e-3, seed=0, log=None,
                 max_samples=None):
    """Capture samples with the plain kernel and train the network on them
    with the GPU kernels (on the first ``max_samples`` samples, in step
    order, when given).  Returns (params, norm, info)."""
    import time
    t0 = time.perf_counter()
    g, n = capture_samples(n_tiles, nsteps)
    if max_samples is not None:
        n = min(n, int(max_samples))
    t1 = time.perf_counter()
    norm = sample_stats(g, n)
    net = MLP(batch=batch, params=init_params(seed), norm=norm)
    xp, yp, _ = g.samples()
    net.use_device_samples(xp, yp, n)
    losses = net.train(epochs=epochs, lr=lr, seed=seed)
    t2 = time.perf_counter()
    params = net.params()
    info = {"samples": int(n), "capture_s": t1 - t0, "train_s": t2 - t1, "batches": int(losses.size),
            "loss_first": float(losses[:16].mean()), "loss_last": float(losses[-16:].mean())}
    if log:
        log(f"[surrogate] {n} samples captured in {t1 - t0:.1f}s; {losses.size} batches in {t2 - t1:.1f}s; "
            f"loss {info['loss_first']:.4f} -> {info['loss_last']:.4f}")
    net.close()
    g.close()
    return params, norm, info


def main(argv=None):
    """python -m cable_amd.surrogate [--tiles N] [--steps K] [--epochs E] [--out F.npz]
    [--hybrid TILES K0 NSTEPS]: capture samples and train on the GPU, then
    optionally run the hybrid step with the trained network on another
    workload of TILES tiles, steps K0..K0+NSTEPS-1 (daylight steps exercise
    the inference); prints the training summary as JSON."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--tiles", type=int, default=262144)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--epochs", type=float, default=2.0)
    ap.add_argument("--max-samples", type=int, default=None)
    ap.add_argument("--out", default=None, help="save params and norm (numpy .npz)")
    ap.add_argument("--hybrid", type=int, nargs=3, metavar=("TILES", "K0", "NSTEPS"), default=None)
    args = ap.parse_args(argv)
    params, norm, info = train_on_gpu(args.tiles, args.steps, args.epochs, max_samples=args.max_samples,
                                      log=lambda m: print(m, file=sys.stderr, flush=True))
    if args.out:
        np.savez(args.out, params=params, norm=norm)
    if args.hybrid:
        from . import gpu
        from .workload import DEFAULT_SEED, Workload
        n, k0, ns = args.hybrid
        wl = Workload(n, k0 + ns, seed=DEFAULT_SEED + 2)
        g = gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False,
                              order=gpu.class_order(wl.iveg, wl.isoilm, wl.vlai))
        g.set_surrogate(params, norm)
        g.run_hybrid(k0, ns, wl.dels, 4, 0.5)
        g.sync()
        fb, ts = g.hybrid_counters()
        info["hybrid"] = {"tiles": n, "first_step": k0, "steps": ns, "device_ms": g.last_run_ms(),
                          "fallback_fraction": fb / ts}
        g.close()
    print(json.dumps(info), flush=True)


if __name__ == "__main__":
    main()
