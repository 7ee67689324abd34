"""Learned CG warm start: per-vertex MLP trained on the solver's own solves.

The reference starts every CG solve from x~ - x (src/mfem/refinement/
kernels.py ``compute_initial_guess``, called at solver.py:530-531) and has no
learned component.  Here a small per-vertex network (16 -> 64 -> 64 -> 3,
GELU, BF16 weights, FP32 accumulation on MFMA; csrc/mfem_mlp.hip) blends
three candidate guesses per vertex into a correction of the previous Newton
step,

    dx0_i = c3_i + s_i (y0 c1_i + y1 c2_i + y2 c3_i) / s_i,
    c1 = x~ - x,  c2 = D^-1 (g - H c1),  c3 = previous Newton step,
    s_i = |c1_i| + |c2_i| + |c3_i|,

from 16 rotation-invariant features (candidate norms and angles, diagonal /
off-diagonal stiffness ratios, mass term, pinning, Newton iteration).  The
PCG starts from whichever of x~ - x (the plain guess), the learned guess and
the previous Newton step has the lowest quadratic-model energy
Phi(x0) = x0.H x0 / 2 - g.x0 (the order of their energy-norm errors, which
bound CG's convergence), so block-Jacobi CG from x~ - x stays the fallback
and the FP64 relative residual the judge of convergence.

Training data are whole solves logged by ``collect_solves`` from plain-PCG
steps (features, candidates, right-hand side, matrix, PCG solution);
``WarmStartNet.fit_solves`` trains on the GPU with the project's
forward/backward kernels, SpMV and Adam on the energy-norm error of the
guess, e^T H e with e = dx0 - dx* (the norm CG's error bound is stated in),
or on the initial residual.  ``validate_cg`` then keeps the trained
correction only if it takes fewer CG iterations over whole steps than the
previous Newton step alone.  (``collect`` / ``fit`` train on per-vertex
(features, PCG solution) pairs with a Euclidean blend loss.)
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _ffi

N_PARAMS = _ffi.MFEM_MLP_NPARAMS
N_IN, N_HID, N_OUT = 16, 64, 16
O_W1 = 0
O_B1 = O_W1 + N_IN * N_HID
O_W2 = O_B1 + N_HID
O_B2 = O_W2 + N_HID * N_HID
O_W3 = O_B2 + N_HID
O_B3 = O_W3 + N_HID * N_OUT
MAX_PARTS = 1024  # workgroups of the gradient kernel (per-workgroup partial gradients)


def _ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _mlp_call(lib, name, *args):
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed ({rc}): {lib.mfem_mlp_last_error().decode()}")


def _blend(C: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """dx0 = c3 + y0 c1 + y1 c2 + y2 c3 for candidates C (n, 3, 3)."""
    return C[:, 2] + y[:, 0:1] * C[:, 0] + y[:, 1:2] * C[:, 1] + y[:, 2:3] * C[:, 2]


def init_params(seed: int = 0) -> np.ndarray:
    """Initial FP32 parameters: scaled normal hidden layers and a zero output
    layer, so the untrained network returns exactly the previous Newton step
    (a random output layer adds per-vertex noise that the CG residual
    amplifies)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_PARAMS, np.float32)
    w[O_W1:O_B1] = rng.normal(0, 1.0 / np.sqrt(N_IN), N_IN * N_HID)
    w[O_W2:O_B2] = rng.normal(0, 1.0 / np.sqrt(N_HID), N_HID * N_HID)
    return w


class WarmStartNet:
    """FP32 master weights, their BF16 copy and Adam state, on one device."""

    def __init__(self, device="cuda", weights: np.ndarray | None = None, seed: int = 0):
        self.lib = _ffi.load()
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        w = init_params(seed) if weights is None else np.asarray(weights, np.float32)
        if w.shape != (N_PARAMS,):
            raise ValueError(f"weights must hold {N_PARAMS} values")
        self.w = torch.as_tensor(w.copy(), device=self.device)
        self.wbf = torch.empty(N_PARAMS, dtype=torch.int16, device=self.device)
        self.adam = torch.zeros(2 * N_PARAMS, dtype=torch.float32, device=self.device)
        self.t = 0
        self._part = None
        self._part_loss = None
        self.refresh()

    def refresh(self) -> None:
        """BF16 copy of the FP32 weights (round to nearest even)."""
        self.wbf.copy_(self.w.to(torch.bfloat16).view(torch.int16))

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """y (n, 3) for BF16 features X (n, 16) stored as int16."""
        X = self._features(X)
        y = torch.empty(X.shape[0], 3, dtype=torch.float32, device=self.device)
        _mlp_call(self.lib, "mfem_mlp_forward", _ptr(X), ctypes.c_int(X.shape[0]), _ptr(self.wbf), _ptr(self.w),
                  _ptr(y), self._stream())
        return y

    def _features(self, X):
        if X.dtype == torch.bfloat16:
            X = X.view(torch.int16)
        if X.dtype != torch.int16 or X.device != self.device or X.dim() != 2 or X.shape[1] != N_IN:
            raise ValueError("features must be a (n, 16) BF16 (or int16 bit-pattern) tensor on the net's device")
        return X.contiguous()

    def _scratch(self):
        if self._part is None:
            self._part = torch.empty(MAX_PARTS * N_PARAMS, dtype=torch.float32, device=self.device)
            self._part_loss = torch.empty(MAX_PARTS, dtype=torch.float32, device=self.device)

    def train_step(self, X, aux, lr: float, loss_out: torch.Tensor | None = None,
                   grad_out: torch.Tensor | None = None) -> None:
        """One Adam step on the mean blend loss of the batch (aux (n, 12) =
        normalised c1, c2, c3 and target).  lr = 0 computes the gradient only
        (into grad_out) and leaves the weights unchanged."""
        X = self._features(X)
        aux = aux.to(device=self.device, dtype=torch.float32).contiguous()
        if aux.shape != (X.shape[0], 12):
            raise ValueError("aux must be (n, 12)")
        self._scratch()
        if lr > 0:
            self.t += 1
        _mlp_call(self.lib, "mfem_mlp_train_step", _ptr(X), _ptr(aux), ctypes.c_int(X.shape[0]), _ptr(self.w),
                  _ptr(self.wbf), _ptr(self.adam), _ptr(self._part), ctypes.c_int(MAX_PARTS), _ptr(self._part_loss),
                  _ptr(loss_out), _ptr(grad_out), ctypes.c_float(lr), ctypes.c_int(max(self.t, 1)), self._stream())

    def backward_step(self, X, dy: torch.Tensor, lr: float, grad_out: torch.Tensor | None = None) -> None:
        """One Adam step from a given output gradient dy = dL/dy (n, 3);
        lr = 0 computes the gradient only (into grad_out)."""
        X = self._features(X)
        dy = dy.to(device=self.device, dtype=torch.float32).contiguous()
        if dy.shape != (X.shape[0], 3):
            raise ValueError("dy must be (n, 3)")
        self._scratch()
        if lr > 0:
            self.t += 1
        _mlp_call(self.lib, "mfem_mlp_backward_step", _ptr(X), _ptr(dy), ctypes.c_int(X.shape[0]), _ptr(self.w),
                  _ptr(self.wbf), _ptr(self.adam), _ptr(self._part), ctypes.c_int(MAX_PARTS), _ptr(grad_out),
                  ctypes.c_float(lr), ctypes.c_int(max(self.t, 1)), self._stream())

    def _solve_loss(self, sv, solver, objective: str, dx0, grad: bool):
        """(loss, dL/ddx0 or None) of one logged solve.  ``residual``:
        log |g - H dx0|^2, gradient -2 H r / |r|^2.  ``energy``:
        log e^T H e with e = dx0 - dx* (the PCG solution of that solve),
        gradient 2 H e / e^T H e."""
        if objective == "residual":
            r = sv["g"] - solver.spmv_with(sv["values"], dx0)
            rr = torch.sum(r.double() * r.double())
            gd = (-2.0 / rr).float() * solver.spmv_with(sv["values"], r) if grad else None
            return torch.log(rr), gd
        if objective == "energy":
            e = dx0 - sv["dx_star"]
            He = solver.spmv_with(sv["values"], e)
            eHe = torch.sum(e.double() * He.double())
            return torch.log(eHe), ((2.0 / eHe).float() * He if grad else None)
        raise ValueError(f"unknown objective {objective!r} (expected 'residual' or 'energy')")

    def objective(self, solves: list, solver, objective: str = "residual", base: bool = False) -> float:
        """Mean of the per-solve objective for the current weights (base:
        for the previous Newton step c3 alone, y = 0)."""
        tot = 0.0
        for sv in solves:
            C = sv["cand"].view(-1, 3, 3)
            dx0 = C[:, 2].contiguous() if base else _blend(C, self.forward(sv["X"]))
            tot += float(self._solve_loss(sv, solver, objective, dx0, False)[0].item())
        return tot / len(solves)

    def residual_objective(self, solves: list, solver) -> float:
        """Mean over solves of log |g - H dx0|^2 for the current weights."""
        return self.objective(solves, solver, "residual")

    def fit_solves(self, solves: list, solver, objective: str = "energy", epochs: int = 20, lr: float = 1e-4,
                   seed: int = 0) -> dict:
        """Trains on logged solves: for each, the objective of the corrected
        guess dx0 = c3 + y0 c1 + y1 c2 + y2 c3 with H that solve's matrix
        (applied by the project's SpMV); dL/dy = c_k . dL/ddx0.  One Adam
        step per solve.  After every epoch the objective is evaluated with
        fixed weights and the best weights are kept, the starting point
        (epoch 0) included, so training never returns a worse guess than it
        started from.  Returns the per-epoch objective, the objective of the
        previous Newton step alone (y = 0) and the epoch kept."""
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = {"objective": objective, "loss": [self.objective(solves, solver, objective)],
                "base": self.objective(solves, solver, objective, base=True), "best_epoch": 0}
        best_w = self.w.clone()
        for ep in range(epochs):
            for i in torch.randperm(len(solves), generator=g).tolist():
                sv = solves[i]
                C = sv["cand"].view(-1, 3, 3)
                dx0 = _blend(C, self.forward(sv["X"]))
                _, gd = self._solve_loss(sv, solver, objective, dx0, True)
                dy = (C * gd[:, None, :]).sum(dim=2)
                self.backward_step(sv["X"], dy, lr)
            hist["loss"].append(self.objective(solves, solver, objective))
            if hist["loss"][-1] < hist["loss"][hist["best_epoch"]]:
                hist["best_epoch"] = ep + 1
                best_w.copy_(self.w)
        self.w.copy_(best_w)
        self.refresh()
        return hist

    def fit_residual(self, solves: list, solver, epochs: int = 20, lr: float = 1e-4, seed: int = 0) -> dict:
        """fit_solves on the initial-residual objective log |g - H dx0|^2."""
        return self.fit_solves(solves, solver, "residual", epochs, lr, seed)

    def validate_cg(self, solver, state, dt: float, steps: int = 2, margin: float = 0.0) -> dict:
        """Model selection on the quantity the guess is for: whole steps from
        ``state`` (the solver's own state restored before each run) with
        these weights and with the network's correction off (zero output
        layer: the previous Newton step, the same fallback); keeps the
        weights only if they take fewer than (1 - margin) times the CG
        iterations of the correction off (zeroing the output layer
        otherwise: margin covers the cost of the network's own launches).
        Returns both counts and the choice."""
        snap = solver.snapshot()
        counts = {}
        trained = self.w.clone()
        zero = trained.clone()
        zero[O_W3:] = 0.0
        try:
            for name, w in (("trained", trained), ("prev_newton_step", zero)):
                self.w.copy_(w)
                self.refresh()
                solver.restore(snap)
                solver.set_warm_start(self)
                a, b = state.clone(), state.clone()
                n = 0
                for k in range(steps):
                    solver.step(a, b, None, None, dt)
                    a, b = b, a
                    n += int(solver.last_stats()[:, 0].sum())
                counts[name] = n
        finally:
            solver.restore(snap)
            solver.set_warm_start(None)
        keep = "trained" if counts["trained"] < (1.0 - margin) * counts["prev_newton_step"] else "prev_newton_step"
        self.w.copy_(trained if keep == "trained" else zero)
        self.refresh()
        return {"cg_iters": counts, "kept": keep, "steps": steps, "margin": margin}

    def loss(self, X, aux) -> float:
        out = torch.zeros(1, dtype=torch.float32, device=self.device)
        self.train_step(X, aux, 0.0, loss_out=out)
        return float(out.item())

    def fit(self, X, aux, epochs: int = 20, batch: int = 65536, lr: float = 2e-3, seed: int = 0,
            val_fraction: float = 0.1) -> dict:
        """Mini-batch Adam on the GPU.  Returns train / validation loss per
        epoch (validation split fixed by ``seed``)."""
        X = self._features(X)
        aux = aux.to(device=self.device, dtype=torch.float32)
        n = X.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        perm = torch.randperm(n, generator=g).to(self.device)
        n_val = int(n * val_fraction)
        val, tr = perm[:n_val], perm[n_val:]
        hist = {"train": [], "val": []}
        loss_buf = torch.zeros(1, dtype=torch.float32, device=self.device)
        for ep in range(epochs):
            order = tr[torch.randperm(len(tr), generator=g).to(self.device)]
            tot, nb = 0.0, 0
            for b0 in range(0, len(order), batch):
                idx = order[b0:b0 + batch]
                if len(idx) < 64:
                    continue
                self.train_step(X[idx], aux[idx], lr, loss_out=loss_buf)
                tot += float(loss_buf.item())
                nb += 1
            hist["train"].append(tot / max(nb, 1))
            if n_val >= 64:
                hist["val"].append(self.loss(X[val], aux[val]))
        return hist

    def weights(self) -> np.ndarray:
        return self.w.cpu().numpy().copy()

    def save(self, path: str) -> None:
        np.save(path, self.weights())

    @classmethod
    def load(cls, path: str, device="cuda") -> "WarmStartNet":
        return cls(device, weights=np.load(path, allow_pickle=False))


def collect(solver, state, dt: float, steps: int, every: int = 1):
    """Runs ``steps`` plain-PCG steps stage by stage on ``state`` (advanced in
    place) and logs, per Newton iteration, the warm-start features, the
    candidates and the PCG solution as training pairs.

    Returns (X (n, 16) int16 BF16 bits, aux (n, 12) float32) on the device,
    aux = (c1, c2, c3, dx - c3) / rms(dx) with rms(dx) the root-mean-square
    solution size of that solve: the loss |sum_k y_k c_k - (dx - c3)|^2 is
    the squared error of the corrected guess relative to the solve's
    solution, so every solve weighs the same (each solve's CG iteration
    count depends on its own relative residual).  Vertices with no scale (at
    rest, all candidates zero) are left out.
    Raises if a device error (grid-barrier timeout) occurred.
    """
    Xs, As = [], []
    out = state.clone()
    b = solver.buffers
    for k in range(steps):
        out.assign(state)
        solver.bind(out)
        solver.begin_step(dt)
        for it in range(solver.iterations):
            solver.update_system(dt)
            solver.assemble(dt)
            solver.ws_features(dt, it)
            solver.global_solve(it, learned=False)
            if k % every == 0:
                s = b["ws_aux"][:, 9:10]
                keep = s[:, 0] > 1e-12 * s.max()
                dx = b["dx"][keep].float()
                rms = torch.sqrt(torch.mean(torch.sum(dx * dx, dim=1))).clamp_min(1e-30)
                cand = b["ws_aux"][keep, :9] * s[keep]
                Xs.append(b["ws_x"][keep].clone())
                As.append(torch.cat([cand / rms, (dx - cand[:, 6:9]) / rms], dim=1))
            solver.local_solve()
            solver.line_search(dt, it)
        solver.end_step(dt)
        state.assign(out)
    if solver.status() != 0:
        raise RuntimeError("device error while logging training pairs (grid barrier timed out)")
    return torch.cat(Xs), torch.cat(As)


def collect_solves(solver, state, dt: float, steps: int) -> list:
    """Like ``collect`` but keeps whole solves for the residual objective:
    per Newton iteration the features X (V, 16), the absolute candidates
    cand (V, 9) = (c1, c2, c3), the right-hand side g (V, 3) and a copy of
    the assembled matrix values (SELL layout, shared pattern), and the
    system's PCG solution dx_star (V, 3)."""
    solves = []
    out = state.clone()
    b = solver.buffers
    for k in range(steps):
        out.assign(state)
        solver.bind(out)
        solver.begin_step(dt)
        for it in range(solver.iterations):
            solver.update_system(dt)
            solver.assemble(dt)
            solver.ws_features(dt, it)
            s = b["ws_aux"][:, 9:10]
            solves.append({"X": b["ws_x"].clone(), "cand": (b["ws_aux"][:, :9] * s).contiguous(),
                           "g": b["g"].clone(), "values": b["sell_val"].clone()})
            solver.global_solve(it, learned=False)
            solves[-1]["dx_star"] = b["dx"].float()  # the PCG solution of this system
            solver.local_solve()
            solver.line_search(dt, it)
        solver.end_step(dt)
        state.assign(out)
    if solver.status() != 0:
        raise RuntimeError("device error while logging solves (grid barrier timed out)")
    return solves
