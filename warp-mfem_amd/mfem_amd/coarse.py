"""Learned smoothing of the PCG's workgroup-local coarse basis.

The reference preconditions its CG with block Jacobi only
(src/mfem/refinement/kernels.py:32-53, solver.py:581-609).  On top of block
Jacobi the persistent PCG can add a workgroup-local rigid-mode correction
(``RefinementSolver.set_coarse_correction``): each PCG workgroup's 256 rows
form one aggregate a whose rigid modes P_a about its centroid, smoothed per
vertex as P'_i = P_i - w_i D_i^-1 (H_aa P)_i, span a coarse space solved
exactly inside the workgroup,

    M^-1 = D^-1 + sum_a P'_a (P'_a^T H_aa P'_a)^+ P'_a^T.

The weights w_i are this module's learned part: a per-vertex network (the
same 16 -> 64 -> 64 -> 3 BF16-MFMA MLP as the warm start, its own weights;
csrc/mfem_mlp.hip) maps 16 features of the assembled system
(``mfem_coarse_features``: the row's coupling outside its aggregate, mass and
contact shares of the diagonal, distance from the aggregate's centroid, ...)
to w_i = clamp(y_0, 0, 2).  Its output layer starts at zero, i.e. w = 0: the
unsmoothed rigid modes, the method without the network.

Training (``CoarseWeightNet.fit``) uses systems the solver logs from its own
steps (``collect_coarse``): for each, the error of the plain start e = dx* -
(x~ - x), and the energy-norm error that k = 5 iterations of the
preconditioned CG leave of it (CG on H y = H e from y = 0: the quantity CG
minimises), log(|e - y_k|_H^2 / |e|_H^2), differentiated through the
iterations.  (Its one-step form, the energy reduction of one optimally
damped preconditioned Richardson step, tracks the CG iteration count over
uniform weights, both smallest near w = 0.75 on the octopus, but rewards
per-vertex variations that CG does not.)  The loss and its gradient with
respect to the network's outputs are formed with torch on the GPU (block
gathers, batched 6 x 6 solves); the network's forward and backward passes
and Adam are the project's kernels.  Model
selection (``validate_cg``) keeps the trained weights only if whole steps
take fewer CG iterations than with the network off (w = 0).
"""
from __future__ import annotations

import numpy as np
import torch

from .learned import O_W3, WarmStartNet

AGG = 256  # rows per aggregate: one PCG workgroup (4 waves x 64 rows)


class BlockSystem:
    """The assembled 3x3-block matrix of one logged solve as gather lists
    (CSR blocks in pattern order) for products with (n, 3, k) tensors, the
    full matrix and its restriction to aggregates (autograd-friendly)."""

    def __init__(self, pattern, n_verts: int, device):
        rowptr = np.asarray(pattern.rowptr, np.int64)
        cols = np.asarray(pattern.cols, np.int64)
        rows = np.repeat(np.arange(n_verts, dtype=np.int64), np.diff(rowptr[: n_verts + 1]))
        lane = np.asarray(pattern.csr_lane, np.int64)[: len(rows)]
        # SELL value index of block b's entry q: ((lane / 64) * 9 + q) * 64 + lane % 64
        q = np.arange(9, dtype=np.int64)
        vidx = ((lane[:, None] // 64) * 9 + q[None, :]) * 64 + (lane[:, None] % 64)
        self.n = n_verts
        self.rows = torch.as_tensor(rows, device=device)
        self.cols = torch.as_tensor(cols[: len(rows)], device=device)
        self.vidx = torch.as_tensor(vidx, device=device)
        self.local = (self.rows // AGG) == (self.cols // AGG)
        self.diag = self.rows == self.cols

    def blocks(self, sell_val: torch.Tensor) -> torch.Tensor:
        """(nnz, 3, 3) block values (float64) from the SELL value array."""
        return sell_val.reshape(-1)[self.vidx].double().view(-1, 3, 3)

    def apply(self, B: torch.Tensor, x: torch.Tensor, local: bool = False) -> torch.Tensor:
        """y = H x (or H_aa x) for x (n, 3, k) with blocks B (nnz, 3, 3)."""
        rows, cols = self.rows, self.cols
        if local:
            m = self.local
            B, rows, cols = B[m], rows[m], cols[m]
        y = torch.zeros_like(x)
        return y.index_add(0, rows, torch.einsum("bij,bjk->bik", B, x[cols]))


def rigid_basis(x: torch.Tensor) -> torch.Tensor:
    """(n, 3, 6) rigid modes of each vertex about its aggregate's centroid,
    positions over the aggregate's RMS radius (as the PCG's set-up forms
    them in FP32): [I | e_k x e]."""
    n = x.shape[0]
    na = -(-n // AGG)
    xp = torch.zeros(na * AGG, 3, dtype=torch.float64, device=x.device)
    xp[:n] = x.double()
    live = torch.zeros(na * AGG, dtype=torch.float64, device=x.device)
    live[:n] = 1.0
    xa, la = xp.view(na, AGG, 3), live.view(na, AGG)
    cnt = la.sum(1, keepdim=True).clamp_min(1.0)
    cen = xa.sum(1) / cnt
    d = (xa - cen[:, None, :]) * la[..., None]
    rad = torch.sqrt((d * d).sum((1, 2)) / cnt[:, 0]).clamp_min(1e-300)
    e = (d / rad[:, None, None]).reshape(-1, 3)[:n].float().double()
    P = torch.zeros(n, 3, 6, dtype=torch.float64, device=x.device)
    P[:, 0, 0] = P[:, 1, 1] = P[:, 2, 2] = 1.0
    P[:, 0, 4], P[:, 0, 5] = e[:, 2], -e[:, 1]
    P[:, 1, 3], P[:, 1, 5] = -e[:, 2], e[:, 0]
    P[:, 2, 3], P[:, 2, 4] = e[:, 1], -e[:, 0]
    return P


class CoarseSystem:
    """Per logged solve: blocks, D^-1, the rigid basis P, Q = D^-1 (H_aa P),
    the test error e and He, ready for the loss of any weights."""

    def __init__(self, bs: BlockSystem, solve: dict):
        self.bs = bs
        n = bs.n
        self.B = bs.blocks(solve["values"])
        D = torch.zeros(n, 3, 3, dtype=torch.float64, device=self.B.device)
        D[bs.rows[bs.diag]] = self.B[bs.diag]
        self.Dinv = torch.linalg.inv(D)
        self.P = rigid_basis(solve["x"][:n])
        HP = bs.apply(self.B, self.P, local=True)
        self.Q = torch.einsum("nij,njk->nik", self.Dinv, HP)
        e = (solve["dx_star"][:n].double() - solve["x0"][:n].double()).view(n, 3, 1)
        self.e = e
        self.He = bs.apply(self.B, e)
        self.eHe = float((e * self.He).sum())

    def apply_M(self, w: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
        """z = M(w)^-1 v for v (n, 3, 1): block Jacobi plus the local coarse
        correction with smoothing weights w (n,)."""
        n = self.bs.n
        na = -(-n // AGG)
        Pp = self.P - w.double()[:, None, None] * self.Q
        HPp = self.bs.apply(self.B, Pp, local=True)
        pad = na * AGG - n

        def blk(t):
            return torch.nn.functional.pad(t.reshape(n * 3, -1), (0, 0, 0, 3 * pad)).view(na, AGG * 3, -1)

        Pa, HPa, va = blk(Pp), blk(HPp), blk(v)
        Ac = Pa.transpose(1, 2) @ HPa
        Ac = 0.5 * (Ac + Ac.transpose(1, 2))
        tr = torch.diagonal(Ac, dim1=1, dim2=2).sum(-1).clamp_min(1e-300)
        Ac = Ac + (1e-12 * tr)[:, None, None] * torch.eye(6, dtype=Ac.dtype, device=Ac.device)
        y = torch.linalg.solve(Ac, Pa.transpose(1, 2) @ va)
        zc = (Pa @ y).reshape(-1, 1)[: n * 3].view(n, 3, 1)
        return torch.einsum("nij,njk->nik", self.Dinv, v) + zc

    def loss(self, w: torch.Tensor, steps: int = 5) -> torch.Tensor:
        """log of the energy-norm error left by ``steps`` iterations of the
        preconditioned CG with weights w on H y = H e from y = 0 (whose
        solution is the plain start's error e), relative to e's: the quantity
        CG minimises, differentiated through the iterations.  steps = 1 is
        one optimally damped preconditioned Richardson step."""
        r = self.He.clone()
        y = torch.zeros_like(r)
        z = self.apply_M(w, r)
        p = z
        rz = (r * z).sum()
        for _ in range(steps):
            Hp = self.bs.apply(self.B, p)
            a = rz / (p * Hp).sum()
            y = y + a * p
            r = r - a * Hp
            z = self.apply_M(w, r)
            rz_new = (r * z).sum()
            p = z + (rz_new / rz) * p
            rz = rz_new
        err = self.e - y
        eH = (err * self.bs.apply(self.B, err)).sum()
        return torch.log((eH / self.eHe).clamp_min(1e-30))


class CoarseWeightNet(WarmStartNet):
    """The coarse-weight network: WarmStartNet's parameters, kernels and
    Adam state, on the coarse features; output 0 is the smoothing weight."""

    def weights_of(self, X: torch.Tensor) -> torch.Tensor:
        return torch.clamp(self.forward(X)[:, 0], 0.0, 2.0)

    def objective(self, systems: list, solves: list, off: bool = False, uniform: float | None = None) -> float:
        """Mean loss over logged systems: the network's weights, w = 0 (off)
        or a uniform weight."""
        tot = 0.0
        with torch.no_grad():
            for cs, sv in zip(systems, solves):
                if off or uniform is not None:
                    w = torch.full((cs.bs.n,), 0.0 if off else float(uniform), device=self.device)
                else:
                    w = self.weights_of(sv["X"][: cs.bs.n])
                tot += float(cs.loss(w))
        return tot / max(len(systems), 1)

    def fit(self, solves: list, pattern, epochs: int = 10, lr: float = 3e-3, seed: int = 0) -> dict:
        """One Adam step per logged system per epoch on the five-step CG loss
        (``CoarseSystem.loss``; dL/dy from torch autograd on the GPU, the
        MLP's backward kernel for the weights); keeps the best epoch by the
        mean loss, epoch 0 (w = 0) included."""
        bs = BlockSystem(pattern, int(solves[0]["n"]), self.device)
        systems = [CoarseSystem(bs, sv) for sv in solves]
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = {"loss": [self.objective(systems, solves)], "best_epoch": 0,
                "off": self.objective(systems, solves, off=True),
                "uniform": self.objective(systems, solves, uniform=0.75)}
        best = self.w.clone()
        for ep in range(epochs):
            for i in torch.randperm(len(systems), generator=g).tolist():
                cs, sv = systems[i], solves[i]
                X = sv["X"][: bs.n]
                y = self.forward(X).detach().requires_grad_(True)
                loss = cs.loss(torch.clamp(y[:, 0], 0.0, 2.0))
                loss.backward()
                self.backward_step(X, y.grad.float(), lr)
            hist["loss"].append(self.objective(systems, solves))
            if hist["loss"][-1] < hist["loss"][hist["best_epoch"]]:
                hist["best_epoch"] = ep + 1
                best.copy_(self.w)
        self.w.copy_(best)
        self.refresh()
        return hist

    def validate_cg(self, solver, state, dt: float, steps: int = 2) -> dict:
        """Whole steps from ``state`` (solver state restored before each)
        with the trained weights and with the network off (zero output
        layer: w = 0); keeps the trained weights only if they take fewer CG
        iterations."""
        snap = solver.snapshot()
        trained = self.w.clone()
        zero = trained.clone()
        zero[O_W3:] = 0.0
        counts = {}
        prev = solver.coarse_net
        try:
            for name, wts in (("trained", trained), ("off", zero)):
                self.w.copy_(wts)
                self.refresh()
                solver.restore(snap)
                solver.set_coarse_net(self)
                a, b = state.clone(), state.clone()
                k = 0
                for _ in range(steps):
                    solver.step(a, b, None, None, dt)
                    a, b = b, a
                    k += int(solver.last_stats()[:, 0].sum())
                counts[name] = k
        finally:
            solver.restore(snap)
            solver.set_coarse_net(prev)
        keep = "trained" if counts["trained"] < counts["off"] else "off"
        self.w.copy_(trained if keep == "trained" else zero)
        self.refresh()
        return {"cg_iters": counts, "kept": keep, "steps": steps}


def collect_coarse(solver, state, dt: float, steps: int) -> list:
    """Plain-PCG steps on ``state`` (advanced in place), logging per Newton
    iteration the coarse features X, the matrix values, positions x, the
    plain start x0 = x~ - x and the PCG solution dx*."""
    solves = []
    out = state.clone()
    b = solver.buffers
    solver._ensure_coarse_buffers()
    n = solver.active_particle_count
    for k in range(steps):
        out.assign(state)
        solver.bind(out)
        solver.begin_step(dt)
        for it in range(solver.iterations):
            solver.update_system(dt)
            solver.assemble(dt)
            solver.coarse_features(dt)
            solves.append({"X": b["cw_x"][:n].clone(), "values": b["sell_val"].clone(),
                           "x": out.particle_q[:n].clone(), "x0": (b["x_tilde"][:n] - out.particle_q[:n]).clone(),
                           "n": n})
            solver.global_solve(it, learned=False)
            solves[-1]["dx_star"] = b["dx"][:n].clone()
            solver.local_solve()
            solver.line_search(dt, it)
        solver.end_step(dt)
        state.assign(out)
    if solver.status() != 0:
        raise RuntimeError("device error while logging systems (grid barrier timed out)")
    return solves
