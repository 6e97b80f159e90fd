"""Differentiable Heun runs: the engine's discrete adjoint behind
torch.autograd.Function.

The state is a float64 CUDA tensor in the engine's device layout (one whole
layer, all blocks; ``Engine.layer_index`` gives each block's offset,
component stride, row-position stride and row order).  ``HeunRun.apply(U0, kn, engine, dts, stride)`` runs
the steps of the frozen dt schedule ``dts`` from U0 (its ghost cells are
recomputed by the ghost fill) and returns the final state; its backward is
the hand-written adjoint of the C ABI (mbm_adj_backward), giving dJ/dU0 on
interior cells and dJ/dKn.
"""
import numpy as np
import torch

from . import L_DATA, NV, Engine, read_deck


class HeunRun(torch.autograd.Function):
    """The engine holds one recorded run at a time. forward keeps U0, Kn, the
    schedule and the number of the recording it made; if another run was
    recorded on the same engine before this graph's backward (a second
    objective, or a no-grad evaluation), backward records this run again
    from the kept U0 before reversing it."""

    @staticmethod
    def _record(engine, U0, kn, dts, stride):
        engine.set_stream(torch.cuda.current_stream().cuda_stream)
        engine.set_kn(float(kn))
        engine.layer_set(L_DATA, U0)
        engine.set_boundary(L_DATA, 0.0)
        engine.adj_forward(tend=float("inf"), cflreg=1.0, dts=dts, stride=stride)
        return engine.adj_runs

    @staticmethod
    def forward(ctx, U0, kn, engine, dts, stride):
        U0c = U0.detach().contiguous()
        run = HeunRun._record(engine, U0c, kn, dts, stride)
        UN = torch.empty_like(U0)
        engine.layer_get(L_DATA, UN)
        ctx.engine, ctx.run, ctx.kn, ctx.dts, ctx.stride = engine, run, float(kn), dts, stride
        ctx.save_for_backward(U0c)
        return UN

    @staticmethod
    def backward(ctx, gUN):
        eng = ctx.engine
        if eng.adj_runs != ctx.run:
            (U0c,) = ctx.saved_tensors
            ctx.run = HeunRun._record(eng, U0c, ctx.kn, ctx.dts, ctx.stride)
        gU0 = torch.empty_like(gUN)
        _, gkn = eng.adj_backward(gUN.contiguous(), gU0)
        return gU0, torch.tensor(gkn, dtype=torch.float64, device=gUN.device), None, None, None


def _block_rows(U, engine):
    """Per block: (rows of the block as (npos, 9, plane) view, row order, nx+6)."""
    out = []
    for (off, plane, pitch, rows), b in zip(engine.layer_index(), engine.blocks):
        npos = int(rows.max()) + 1
        idx = torch.as_tensor(rows, device=U.device)
        out.append((U[off:off + npos * pitch].view(npos, NV, plane), idx, b["nx"] + 6))
    return out


def block_views(U, engine):
    """(9, ny+6, nx+6) cells of each block of a layer-layout tensor (gathered
    through the row order; differentiable)."""
    return [blk.index_select(0, idx)[:, :, :lx].permute(1, 0, 2)
            for blk, idx, lx in _block_rows(U, engine)]


def set_interiors(U, engine, values):
    """Writes (9, ny, nx) interior values of every block into U in place
    (differentiable; U must not be a leaf that requires grad)."""
    for (blk, idx, lx), v in zip(_block_rows(U, engine), values):
        rows = blk.index_select(0, idx)
        rows[3:-3, :, 3:lx - 3] = v.permute(1, 0, 2)
        blk.index_copy_(0, idx, rows)


def theta(Ub):
    """theta = (2E/rho - |v|^2)/3 of a (9, ...) block view."""
    rho = Ub[0]
    v1, v2 = Ub[1] / rho, Ub[2] / rho
    return (2 * Ub[3] / rho - v1 * v1 - v2 * v2) / 3


def state_from_theta(th, rho, extra=None):
    """Conservative state with temperature theta. extra: None (at rest,
    T_ij = theta delta_ij, q = 0) or fields (rho~, v1, v2, dT11, dT22, T12, q1, q2)
    added to the state at rest."""
    z = torch.zeros_like(th)
    if extra is None:
        p = rho * th
        return torch.stack([rho + z, z, z, 1.5 * p, p, z, p, z, z])
    dr, v1, v2, d11, d22, t12, q1, q2 = extra
    r = rho + dr
    t11, t22 = th + d11, th + d22
    E = r * (0.5 * (v1 * v1 + v2 * v2) + 1.5 * th)
    return torch.stack([r + z, r * v1, r * v2, E, r * (t11 + v1 * v1), r * (t12 + v1 * v2),
                        r * (t22 + v2 * v2), q1 + E * v1 + r * (t11 * v1 + t12 * v2),
                        q2 + E * v2 + r * (t12 * v1 + t22 * v2)])


def smooth_fields(X, Y, rng, amplitude):
    """Eight smooth random fields (sums of low Fourier modes) on cell centres."""
    out = []
    for scale in (0.1, 0.1, 0.1, 0.05, 0.05, 0.05, 0.02, 0.02):
        f = np.zeros_like(X)
        for _ in range(3):
            kx, ky, ph = rng.uniform(0.5, 3.0), rng.uniform(0.5, 3.0), rng.uniform(0, 2 * np.pi)
            f += np.sin(kx * X + ky * Y + ph)
        out.append(amplitude * scale * f / 3)
    return out


class TemperatureInversion:
    """J(theta0, Kn) = sum over interior cells of (theta(t_end) - theta_target)^2
    for a deck started from a temperature field theta0 at uniform density."""

    def __init__(self, deck_path, kn=None, rho=1.0, seed=0, device="cuda", generic=0.0,
                 limiter=None):
        """generic > 0 adds smooth fields of that relative amplitude to every
        primitive of the base state (theta0 included), so that no slope is
        exactly zero (at a state at rest, or where the tanh profile is flat, the
        limiters sit on their kinks and finite differences measure the
        limiter's nonlinear response, not a derivative)."""
        self.deck = read_deck(deck_path)
        self.kn = self.deck.kn if kn is None else kn
        self.rho = rho
        kw = {} if limiter is None else {"limiter": limiter}
        self.engine = Engine.from_deck(self.deck, scenario="tempjump", **kw)
        self.engine.set_kn(self.kn)
        self.device = device
        self.engine.init()
        U = torch.empty(self.engine.layer_elems(), dtype=torch.float64, device=device)
        self.engine.layer_get(L_DATA, U)
        self.U_init = U
        self.theta0 = [theta(v)[3:-3, 3:-3].clone() for v in block_views(U, self.engine)]
        rng = np.random.default_rng(seed)
        self.target = []
        self.extra = []
        for k, b in enumerate(self.deck.blocks):
            bx, by, dx, dy = self.deck.geom(k)
            x = bx + dx * np.arange(3, b["nx"] + 3)
            y = by + dy * np.arange(3, b["ny"] + 3)
            X, Y = np.meshgrid(x, y)
            cx, cy = rng.uniform(-0.3, 0.3, size=2)
            amp, wid = rng.uniform(0.5, 1.5), rng.uniform(0.2, 0.4)
            t = 1.0 + amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * wid ** 2))
            self.target.append(torch.tensor(t, dtype=torch.float64, device=device))
            if generic > 0:
                self.extra.append([torch.tensor(f, dtype=torch.float64, device=device)
                                   for f in smooth_fields(X, Y, rng, generic)])
                # theta itself: the tanh profile is exactly flat away from the jump
                dth = smooth_fields(X, Y, rng, generic)[3]
                self.theta0[k] = self.theta0[k] + torch.tensor(dth, dtype=torch.float64,
                                                               device=device)
            else:
                self.extra.append(None)
        self.dts = None

    def record_schedule(self, tend, maxsteps=0):
        """Runs the adaptive controller once from the base state and keeps its dt."""
        self.engine.layer_set(L_DATA, self.initial_state(self.theta0))
        self.engine.set_boundary(L_DATA, 0.0)
        n, t = self.engine.adj_forward(tend=tend, cflreg=self.deck.cfl, maxsteps=maxsteps)
        self.dts = self.engine.adj_schedule()
        return n, t

    def initial_state(self, theta0):
        U = self.U_init.clone()
        set_interiors(U, self.engine, [state_from_theta(th, self.rho, ex)
                                       for th, ex in zip(theta0, self.extra)])
        return U

    def final_theta(self, theta0, kn):
        """Interior theta of every block at the end of the run (no gradient)."""
        with torch.no_grad():
            UN = HeunRun.apply(self.initial_state(theta0), kn, self.engine, self.dts, None)
            return [theta(v)[3:-3, 3:-3].clone() for v in block_views(UN, self.engine)]

    def objective_difference(self, th_plus, th_minus):
        """J(+) - J(-) from two final temperature fields, summed cell by cell as
        (th+ - th-)(th+ + th- - 2 target): the difference of two nearly equal
        large sums is never formed."""
        return float(sum(((p - m) * (p + m - 2 * t)).sum()
                         for p, m, t in zip(th_plus, th_minus, self.target)))

    def objective(self, theta0, kn, stride=None):
        U0 = self.initial_state(theta0)
        UN = HeunRun.apply(U0, kn, self.engine, self.dts, stride)
        J = 0.0
        for v, tgt in zip(block_views(UN, self.engine), self.target):
            J = J + ((theta(v)[3:-3, 3:-3] - tgt) ** 2).sum()
        return J
