"""The learned warm start: a small MLP trained on the plain solver's results, run by this package's HIP kernels.

The network (kernels/mlp.hip) maps a cell's call inputs to what the call will do:
- inputs (120 for NCHO): log10 x clipped at -30, ln T, ln P, ln dt and log10 J with 0 mapped to -20;
- outputs (71): ln of the step size the call ends with, ln(1 + accepted steps) and the change of log10 x.

Inputs and outputs are standardised with the training set's mean and spread. The network has 4 hidden layers of
width 256 with GELU. Weights and activations are BF16, products accumulate in FP32 on the MFMA units, and training
keeps an FP32 master copy updated by Adam. Everything runs on the GPU through the C ABI's launch; PyTorch only holds
the memory.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _hip, _kernels
from ._constants import COUNT_SLOTS, H_MIN
from .kincuda import CudaSolver

TILE = 64
INNER = 32
J_ZERO = -20.0  # log10 J of a dark branch
X_FLOOR = -30.0  # log10 x is clipped here


def pad(n: int, to: int = TILE) -> int:
    return -(-n // to) * to


_module: dict[int, _hip.Module] = {}


def module(device: int) -> _hip.Module:
    """The MLP kernels (not network-specific), compiled once and loaded once per device."""
    if device not in _module:
        _module[device] = _hip.Module(_hip.compile_source(_kernels.mlp_source(), "mlp.hip"), device)
    return _module[device]


@dataclass
class Layer:
    w: torch.Tensor  # FP32 [n_out_pad, n_in_pad], the master copy
    b: torch.Tensor  # FP32 [n_out_pad]
    wb: torch.Tensor  # BF16 copy of w (as int16 storage)
    mw: torch.Tensor
    vw: torch.Tensor
    mb: torch.Tensor
    vb: torch.Tensor
    gw: torch.Tensor
    gb: torch.Tensor


class MLP:
    """An MLP of n_in -> width x depth (GELU) -> n_out on one device, padded to the kernels' tiles."""

    def __init__(self, n_in: int, n_out: int, device, width: int = 256, depth: int = 4, seed: int = 0) -> None:
        self.dev = torch.device(device)
        self.n_in, self.n_out, self.width, self.depth = n_in, n_out, width, depth
        self.dims = [pad(n_in)] + [pad(width)] * depth + [pad(n_out)]
        self.true_dims = [n_in] + [width] * depth + [n_out]
        rng = np.random.default_rng(seed)
        self.layers: list[Layer] = []
        for l in range(depth + 1):
            k, n, kt, nt = self.dims[l], self.dims[l + 1], self.true_dims[l], self.true_dims[l + 1]
            w = np.zeros((n, k), np.float32)
            w[:nt, :kt] = rng.standard_normal((nt, kt)) * math.sqrt(2.0 / kt)  # He init
            if l == depth:
                w *= 0.1
            self.layers.append(self._layer(w, np.zeros(n, np.float32)))
        self.t = 0
        self._mod = module(self.dev.index if self.dev.index is not None else torch.cuda.current_device())
        self._fn = {name: self._mod.function(name) for name in (
            "mlp_gemm_f32", "mlp_gemm_bias_gelu", "mlp_gemm_bias", "mlp_loss", "mlp_gelu_backward",
            "mlp_column_sum", "mlp_column_total", "mlp_adam", "mlp_to_bf16")}  # fmt: skip
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self._sync_bf16()

    def _layer(self, w: np.ndarray, b: np.ndarray) -> Layer:
        d = self.dev
        wt = torch.from_numpy(w).to(d)
        bt = torch.from_numpy(b).to(d)
        z = torch.zeros_like
        return Layer(wt, bt, torch.zeros(w.shape, dtype=torch.int16, device=d), z(wt), z(wt), z(bt), z(bt), z(wt),
                     z(bt))  # fmt: skip

    # ---- launches ----
    def _stream(self) -> int:
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _elementwise(self, name: str, n: int, args: list) -> None:
        _hip.launch(self._fn[name], max(1, -(-n // 256)), 256, 0, self._stream(), args)

    def _gemm(self, name: str, M: int, N: int, K: int, A, sam, sak, B, sbk, sbn, tail: list) -> None:
        assert M % TILE == 0 and N % TILE == 0 and K % INNER == 0, (M, N, K)
        p, L = _hip.ptr, ctypes.c_long
        args = [p(A), L(sam), L(sak), p(B), L(sbk), L(sbn), ctypes.c_int(K), ctypes.c_int(N // TILE), *tail]
        _hip.launch(self._fn[name], (N // TILE) * (M // TILE), 256, 0, self._stream(), args)

    def _sync_bf16(self) -> None:
        for ly in self.layers:
            self._elementwise("mlp_to_bf16", ly.w.numel(), [_hip.ptr(ly.w), _hip.ptr(ly.wb), ctypes.c_long(ly.w.numel())])

    # ---- forward and backward ----
    def forward(self, X: torch.Tensor, keep: bool = False) -> torch.Tensor:
        """Y [M, n_out_pad] FP32 of the BF16 inputs X [M, n_in_pad] (int16 storage); M a multiple of 64."""
        M = X.shape[0]
        p = _hip.ptr
        a = X
        acts, pre = [X], []
        for l, ly in enumerate(self.layers):
            K, N = self.dims[l], self.dims[l + 1]
            z = torch.empty(M, N, dtype=torch.float32, device=self.dev)
            if l < self.depth:
                nxt = torch.empty(M, N, dtype=torch.int16, device=self.dev)
                self._gemm("mlp_gemm_bias_gelu", M, N, K, a, K, 1, ly.wb, 1, K,
                           [p(ly.b), p(z), p(nxt), ctypes.c_long(N)])  # fmt: skip
                pre.append(z)
                acts.append(nxt)
                a = nxt
            else:
                self._gemm("mlp_gemm_bias", M, N, K, a, K, 1, ly.wb, 1, K, [p(ly.b), p(z), ctypes.c_long(N)])
                y = z
        if keep:
            self._acts, self._pre = acts, pre
        return y

    def backward(self, dY: torch.Tensor) -> None:
        """Gradients of every layer (gw, gb) from dY [M, n_out_pad] (BF16) after forward(keep=True)."""
        M = dY.shape[0]
        p = _hip.ptr
        dz = dY
        for l in range(self.depth, -1, -1):
            ly = self.layers[l]
            K, N = self.dims[l], self.dims[l + 1]
            x = self._acts[l]
            # dW = dZ^T X: (n, m) of dZ at n + m N; (m, k) of X at m K + k
            self._gemm("mlp_gemm_f32", N, K, M, dz, 1, N, x, K, 1, [p(ly.gw), ctypes.c_long(K), ctypes.c_float(1.0)])
            # the bias gradient: column sums of dZ over blocks of rows, then over the blocks
            rows_per_block = 256
            rb = -(-M // rows_per_block)
            partial = torch.empty(rb, N, dtype=torch.float32, device=self.dev)
            _hip.launch(self._fn["mlp_column_sum"], (N // TILE) * rb, 256, 0, self._stream(),
                        [p(dz), ctypes.c_int(M), ctypes.c_int(N), ctypes.c_int(rows_per_block), ctypes.c_int(N // TILE),
                         p(partial)])  # fmt: skip
            self._elementwise("mlp_column_total", N, [p(partial), ctypes.c_int(rb), ctypes.c_int(N), p(ly.gb)])
            if l > 0:
                da = torch.empty(M, K, dtype=torch.float32, device=self.dev)
                # dA = dZ W: (m, n) of dZ at m N + n; (n, k) of W at n K + k
                self._gemm("mlp_gemm_f32", M, K, N, dz, N, 1, ly.wb, K, 1, [p(da), ctypes.c_long(K), ctypes.c_float(1.0)])
                nz = torch.empty(M, K, dtype=torch.int16, device=self.dev)
                self._elementwise("mlp_gelu_backward", M * K, [p(da), p(self._pre[l - 1]), p(nz), ctypes.c_long(M * K)])
                dz = nz

    def loss_grad(self, Y: torch.Tensor, T: torch.Tensor, w: torch.Tensor, rows: int) -> torch.Tensor:
        """dY (BF16) of the weighted mean squared error over the first rows rows; adds the sum of squares to loss."""
        M, ldy = Y.shape
        dY = torch.empty(M, ldy, dtype=torch.int16, device=self.dev)
        p = _hip.ptr
        self._elementwise("mlp_loss", M * ldy, [p(Y), p(T), p(w), ctypes.c_int(rows), ctypes.c_int(self.n_out),
                                                ctypes.c_long(ldy), ctypes.c_long(M * ldy), p(dY), p(self.loss)])  # fmt: skip
        return dY

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, decay: float = 0.0) -> None:
        self.t += 1
        bc1, bc2 = 1.0 - beta1**self.t, 1.0 - beta2**self.t
        p, f = _hip.ptr, ctypes.c_float
        for ly in self.layers:
            for prm, m, v, g, pb in ((ly.w, ly.mw, ly.vw, ly.gw, ly.wb), (ly.b, ly.mb, ly.vb, ly.gb, None)):
                n = prm.numel()
                self._elementwise("mlp_adam", n, [p(prm), p(m), p(v), p(g), p(pb), ctypes.c_long(n), f(lr), f(beta1),
                                                  f(beta2), f(eps), f(bc1), f(bc2), f(decay)])  # fmt: skip

    def to_bf16(self, x: torch.Tensor) -> torch.Tensor:
        """FP32 -> BF16 (int16 storage) by the kernels' rounding."""
        y = torch.empty(x.shape, dtype=torch.int16, device=self.dev)
        self._elementwise("mlp_to_bf16", x.numel(), [_hip.ptr(x), _hip.ptr(y), ctypes.c_long(x.numel())])
        return y

    def state(self) -> dict[str, np.ndarray]:
        out = {}
        for l, ly in enumerate(self.layers):
            out[f"w{l}"] = ly.w.cpu().numpy()
            out[f"b{l}"] = ly.b.cpu().numpy()
        return out

    def load(self, state: dict[str, np.ndarray]) -> None:
        for l, ly in enumerate(self.layers):
            ly.w.copy_(torch.from_numpy(np.asarray(state[f"w{l}"], np.float32)))
            ly.b.copy_(torch.from_numpy(np.asarray(state[f"b{l}"], np.float32)))
        self._sync_bf16()


# ---- the surrogate of one network's call ----


class Surrogate:
    """The MLP of one network with its standardisation: features and targets of calls, training and predictions.

    Features of a cell: log10 max(x, 1e-30) per species, ln T, ln P, ln dt, log10 J per branch (J = 0 as -20).
    Targets of a call: ln h_out, ln(1 + accepted steps), log10 max(x_out, 1e-30) - log10 max(x_in, 1e-30).
    """

    def __init__(self, net, device, width: int = 256, depth: int = 4, seed: int = 0) -> None:
        self.dev = torch.device(device)
        self.species = list(net.species)
        self.ns, self.nj = len(net.species), len(net.photo)
        self.n_in, self.n_out = self.ns + 3 + self.nj, self.ns + 2
        self.mlp = MLP(self.n_in, self.n_out, self.dev, width, depth, seed)
        self.atoms = torch.from_numpy(np.ascontiguousarray(net.atoms, np.float64)).to(self.dev)
        self.nel = int(net.atoms.shape[1])
        self.mu_in = torch.zeros(self.n_in, dtype=torch.float32, device=self.dev)
        self.sd_in = torch.ones_like(self.mu_in)
        self.mu_out = torch.zeros(self.n_out, dtype=torch.float32, device=self.dev)
        self.sd_out = torch.ones_like(self.mu_out)
        self.w_out = torch.ones(self.n_out, dtype=torch.float32, device=self.dev)
        self._fn = {n: self.mlp._mod.function(n) for n in ("mlp_features", "mlp_targets", "mlp_decode")}
        self.history: list[float] = []

    # features and targets, in float64 on the host side only for the standardisation's statistics
    def raw_features(self, x, T, P, dt, J) -> torch.Tensor:
        cols = [torch.log10(torch.clamp(x, min=1e-30)), torch.log(T)[:, None], torch.log(P)[:, None],
                torch.log(torch.clamp(dt, min=1e-30))[:, None]]  # fmt: skip
        if self.nj:
            Jz = torch.zeros(x.shape[0], self.nj, dtype=x.dtype, device=x.device) if J is None else J
            cols.append(torch.where(Jz > 0, torch.log10(torch.clamp(Jz, min=1e-300)), torch.full_like(Jz, J_ZERO)))
        return torch.cat(cols, 1)

    def raw_targets(self, x0, x1, h, steps) -> torch.Tensor:
        return torch.cat([torch.log(torch.clamp(h, min=1e-30))[:, None], torch.log1p(torch.clamp(steps, min=0))[:, None],
                          torch.log10(torch.clamp(x1, min=1e-30)) - torch.log10(torch.clamp(x0, min=1e-30))], 1)  # fmt: skip

    def standardise(self, data: dict, sample: int = 262144, seed: int = 0) -> None:
        n = data["x0"].shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        idx = torch.randperm(n, generator=g)[: min(n, sample)].to(self.dev)
        f = self.raw_features(data["x0"][idx], data["T"][idx], data["P"][idx], data["dt"][idx],
                              None if data["J"] is None else data["J"][idx])  # fmt: skip
        t = self.raw_targets(data["x0"][idx], data["x1"][idx], data["h1"][idx], data["steps"][idx])
        self.mu_in.copy_(f.mean(0).float())
        self.sd_in.copy_(torch.clamp(f.std(0), min=1e-3).float())
        self.mu_out.copy_(t.mean(0).float())
        self.sd_out.copy_(torch.clamp(t.std(0), min=1e-3).float())

    def features(self, x, T, P, dt, J, rows: int) -> torch.Tensor:
        """BF16 features [rows, n_in padded] of the cells (rows >= cells; the rest 0), by the feature kernel."""
        B = x.shape[0]
        ldf = self.mlp.dims[0]
        F = torch.empty(rows, ldf, dtype=torch.int16, device=self.dev)
        p = _hip.ptr
        n = rows * ldf
        _hip.launch(self._fn["mlp_features"], -(-n // 256), 256, 0, self.mlp._stream(),
                    [p(x), p(T), p(P), p(dt), p(J), ctypes.c_int(B), ctypes.c_int(self.ns), ctypes.c_int(self.nj),
                     p(self.mu_in), p(self.sd_in), p(F), ctypes.c_int(rows), ctypes.c_int(ldf)])  # fmt: skip
        return F

    def targets(self, x0, x1, h, steps) -> torch.Tensor:
        B = x0.shape[0]
        Tg = torch.empty(B, self.n_out, dtype=torch.float32, device=self.dev)
        p = _hip.ptr
        _hip.launch(self._fn["mlp_targets"], max(1, -(-(B * self.n_out) // 256)), 256, 0, self.mlp._stream(),
                    [p(x0), p(x1), p(h), p(steps), ctypes.c_int(B), ctypes.c_int(self.ns), p(self.mu_out),
                     p(self.sd_out), p(Tg)])  # fmt: skip
        return Tg

    # ---- training ----
    def fit(self, data: dict, epochs: float = 4.0, batch: int = 8192, lr: float = 2e-3, seed: int = 0,
            log=None) -> list[float]:
        """Adam on the data's calls (dict of device tensors: x0, T, P, dt, J, x1, h1, steps); the mean loss of each
        epoch (standardised units)."""
        n = data["x0"].shape[0]
        batch = pad(min(batch, pad(n)))
        steps = max(1, int(epochs * n / batch))
        g = torch.Generator(device="cpu").manual_seed(seed)
        perm = torch.randperm(n, generator=g).to(self.dev)
        pos = 0
        epoch_loss, seen, out = 0.0, 0, []
        self.mlp.loss.zero_()
        for it in range(steps):
            if pos + batch > n:
                perm = torch.randperm(n, generator=g).to(self.dev)
                pos = 0
            idx = perm[pos : pos + batch]
            pos += batch
            J = None if data["J"] is None else data["J"][idx]
            X = self.features(data["x0"][idx], data["T"][idx], data["P"][idx], data["dt"][idx], J, batch)
            Tg = self.targets(data["x0"][idx], data["x1"][idx], data["h1"][idx], data["steps"][idx])
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * it / steps))  # cosine decay
            Y = self.mlp.forward(X, keep=True)
            self.mlp.backward(self.mlp.loss_grad(Y, Tg, self.w_out, batch))
            self.mlp.adam(rate)
            seen += batch
            if seen >= n or it == steps - 1:
                epoch_loss = float(self.mlp.loss) / (seen * self.n_out)
                out.append(epoch_loss)
                if log:
                    log(f"surrogate: step {it + 1}/{steps}, mean loss {epoch_loss:.4f}")
                self.mlp.loss.zero_()
                seen = 0
        self.history += out
        return out

    # ---- predictions ----
    def predict(self, x, T, P, dt, J, hmin: float, proposals: bool = True):
        """(h0 [B], cost [B], proposals [B, NS] or None) of the cells: h0 the warm-start step, cost the expected
        accepted steps, a proposal's first entry NaN where its element shares could not be matched; proposals=False
        leaves them out (and their projection onto the element shares, most of the decoding's work)."""
        B = x.shape[0]
        rows = pad(B)
        X = self.features(x, T, P, dt, J, rows)
        Y = self.mlp.forward(X)
        h0 = torch.empty(B, dtype=torch.float64, device=self.dev)
        cost = torch.empty_like(h0)
        xp = torch.empty(B, self.ns, dtype=torch.float64, device=self.dev) if proposals else None
        p = _hip.ptr
        _hip.launch(self._fn["mlp_decode"], max(1, -(-B // 256)), 256, 0, self.mlp._stream(),
                    [p(Y), ctypes.c_int(Y.shape[1]), p(self.mu_out), p(self.sd_out), p(x), p(dt), p(self.atoms),
                     ctypes.c_int(B), ctypes.c_int(self.ns), ctypes.c_int(self.nel), ctypes.c_double(hmin), p(h0),
                     p(cost), p(xp)])  # fmt: skip
        return h0, cost, xp

    def evaluate(self, data: dict, rows: int = 65536) -> dict[str, float]:
        """Mean absolute errors in the targets' own units on up to rows calls."""
        n = min(rows, data["x0"].shape[0])
        sl = slice(0, n)
        J = None if data["J"] is None else data["J"][sl]
        X = self.features(data["x0"][sl], data["T"][sl], data["P"][sl], data["dt"][sl], J, pad(n))
        Y = self.mlp.forward(X)[:n, : self.n_out] * self.sd_out + self.mu_out
        t = self.raw_targets(data["x0"][sl], data["x1"][sl], data["h1"][sl], data["steps"][sl]).float()
        e = (Y - t).abs()
        return {"ln_h": float(e[:, 0].mean()), "ln_steps": float(e[:, 1].mean()), "dlog10x": float(e[:, 2:].mean())}

    def save(self, path) -> None:
        arrays = self.mlp.state()
        for k in ("mu_in", "sd_in", "mu_out", "sd_out"):
            arrays[k] = getattr(self, k).cpu().numpy()
        arrays["species"] = np.array(self.species)
        np.savez(path, **arrays)

    def load(self, path) -> None:
        with np.load(path, allow_pickle=False) as z:
            if list(z["species"]) != self.species:
                raise ValueError(f"{path}: a surrogate of another network")
            self.mlp.load({k: z[k] for k in z.files if k[0] in "wb" and k[1:].isdigit()})
            for k in ("mu_in", "sd_in", "mu_out", "sd_out"):
                getattr(self, k).copy_(torch.from_numpy(z[k]))


def generate(solver, sequences: list[dict], log=None) -> dict:
    """Calls of the plain solver and what they did, as training data (device tensors).

    Each sequence is a dict of numpy inputs (x, T, P, h, J and a list dts of per-call dt arrays, each with its J or
    None for dark): the calls follow one another, each starting from the previous one's x and h."""
    dev = solver.dev
    keys = ("x0", "T", "P", "dt", "J", "x1", "h1", "steps")
    parts: dict[str, list] = {k: [] for k in keys}

    def g(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    for sq in sequences:
        x, h = g(sq["x"]), g(sq["h"])
        T, P = g(sq["T"]), g(sq["P"])
        for dt_np, J_np in sq["calls"]:
            dt = g(dt_np)
            J = None if J_np is None else g(J_np)
            r = solver(x, T, P, dt, h, J=J)
            keep = r.done
            for k, v in (("x0", x), ("T", T), ("P", P), ("dt", dt), ("x1", r.x), ("h1", r.h), ("steps", r.steps)):
                parts[k].append(v[keep])
            parts["J"].append(None if J is None else J[keep])
            x, h = r.x, r.h
            if log:
                log(f"data: {int(keep.sum())} calls of dt median {float(dt.median()):.3g} s")
    out = {k: torch.cat(v) for k, v in parts.items() if k != "J"}
    if any(j is None for j in parts["J"]):
        nj = next((j.shape[1] for j in parts["J"] if j is not None), 0)
        parts["J"] = [torch.zeros(len(xx), nj, dtype=torch.float64, device=dev) if j is None else j
                      for j, xx in zip(parts["J"], parts["x0"])]  # fmt: skip
    out["J"] = torch.cat(parts["J"]) if parts["J"][0] is not None else None
    return out


class LearnedSolver(CudaSolver):
    """CudaSolver with the learned warm start of a trained Surrogate.

    The surrogate's cost estimate orders the queue (costliest first) where no previous call over the same cells
    measured it. Its step size starts the cells that come
    cold (warm="cold": a caller's h that is not finite or below COLD times dt, as a first call's is; warm="all":
    every cell; warm="none"). Its proposed end state is tried first when propose is set: it is kept only when one
    error-controlled try of the whole dt from it passes and moves it by no more than the tolerance. Every other cell
    is solved by the plain method from its input, and the guard and the recovery judge every result as in the plain
    form. The call, results, counts and refusals are CudaSolver's.

    Proposals are off by default: a try from a state tests that the state is stationary over dt, not that it is the
    cell's state at the end of dt, and where chemistry is slow a wrong proposal passes (measured on the NCHO
    workload: species off by several times).

    A call with warm="cold", no proposals and the previous call's tries to order the queue runs the inference on its
    cold cells' rows only (none if no cell is cold); predicted_cost is then None.
    """

    COLD = 1e-6

    def __init__(self, net, device, surrogate: Surrogate, *, drift=None, photo=False, propose: bool = False,
                 warm: str = "cold", precision: str = "fp64") -> None:  # fmt: skip
        if warm not in ("cold", "all", "none"):
            raise ValueError(f"warm: 'cold', 'all' or 'none' expected, got {warm!r}")
        super().__init__(net, device, drift=drift, photo=photo, precision=precision)
        self.surrogate, self.propose, self.warm = surrogate, propose, warm
        self.predicted_cost: torch.Tensor | None = None

    def _plan(self, x, T, P, dt, h_start, jx, B):
        # a non-finite h the warm start replaces is counted as the call kernel counts the ones it sees
        planned = torch.zeros_like(self.count)
        if self.warm != "none":
            planned[COUNT_SLOTS.index("nonfinite_h")] = (~torch.isfinite(h_start)).sum()
        prev = self.tries
        measured = prev is not None and prev.shape[0] == B
        if self.warm == "cold" and not self.propose and measured:
            # only the cold cells' first steps come from the prediction (the queue follows the measured tries):
            # inference on their rows alone, which give the same values as in a whole batch (one host sync)
            cold = ~torch.isfinite(h_start) | (h_start < self.COLD * dt)
            idx = torch.nonzero(cold).squeeze(1)
            if idx.numel():
                h0, _, _ = self.surrogate.predict(x[idx], T[idx], P[idx], dt[idx], None if jx is None else jx[idx],
                                                  H_MIN, proposals=False)  # fmt: skip
                h_start[idx] = h0
            self.predicted_cost = None
            return torch.argsort(prev, descending=True).to(torch.int32), None, planned
        h0, cost, xp = self.surrogate.predict(x, T, P, dt, jx, H_MIN, proposals=self.propose)
        if self.warm == "all":
            h_start.copy_(h0)
        elif self.warm == "cold":
            cold = ~torch.isfinite(h_start) | (h_start < self.COLD * dt)
            h_start.copy_(torch.where(cold, h0, h_start))
        self.predicted_cost = cost
        # the queue: the previous call's tries when it had these cells (a measured cost), else the predicted one
        prev = self.tries
        key = prev if prev is not None and prev.shape[0] == B else cost
        order = torch.argsort(key, descending=True).to(torch.int32)
        return order, (xp if self.propose else None), planned
