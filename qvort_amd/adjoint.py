"""Multi-step sensitivities of filament objectives through the GPU velocity.

The velocity u = U(x) of the direct sum (velocity='BS': LIA term +
biot_savart + biot_savart_shift, timestep.f90:218-259, 462-507) is a
``torch.autograd.Function`` whose backward is the library's hand-written
vector-Jacobian product (qv_vjp_device); the position update is the
reference's Adams-Bashforth ladder (timestep.f90:48-71) and periodic wrap
(periodic.f90:188-213) written in torch operations, so autograd carries
gradients through many steps.  Rollouts are checkpointed: every segment of
``checkpoint_every`` steps keeps only its inputs and is recomputed in the
backward pass.

Objectives: the total line length (line.f90:43) and the energy of
``energy_info`` (diagnostic.f90:259-278), E = Gamma sum_i l_i/2 [u_i . (x_i x
s'_i) + u_{i+1} . (x_{i+1} x s'_{i+1})]; the reference notes that the latter
is not an energy under periodic boundaries, it is still a smooth function of
the positions and serves as the second objective.
"""
from __future__ import annotations

import math

import torch
from torch.utils.checkpoint import checkpoint

from . import gpu


class Engine:
    """A device context bound to torch's current stream, holding the links."""

    def __init__(self, fil, *, box_size: float, boundary: str = "periodic",
                 quant_circ: float = 9.97e-4, corea: float = 1.3e-8, deriv_order: int = 2,
                 precision: str = "fp64"):
        if not torch.cuda.is_available():
            raise RuntimeError("the adjoint harness runs on the GPU only")
        self.params = gpu.params(velocity="BS", boundary=boundary, box_size=box_size,
                                 quant_circ=quant_circ, corea=corea, deriv_order=deriv_order,
                                 precision=precision)
        self.ctx = gpu.Context(self.params)
        self.ctx.upload(fil.x, fil.infront, fil.behind)
        self.n = fil.pcount
        self.box_size = box_size
        self.boundary = boundary
        self.quant_circ = quant_circ
        dev = torch.device("cuda")
        self.infront = torch.as_tensor(fil.infront.astype("int64") - 1, device=dev)
        self.behind = torch.as_tensor(fil.behind.astype("int64") - 1, device=dev)
        self.active = torch.as_tensor(fil.infront != 0, device=dev)

    def _bind(self):
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)

    def close(self):
        self.ctx.close()


class _Velocity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, eng):
        x = x.contiguous()
        u = torch.empty_like(x)
        eng._bind()
        eng.ctx.velocity_device(x.data_ptr(), u.data_ptr())
        ctx.save_for_backward(x)
        ctx.eng = eng
        return u

    @staticmethod
    def backward(ctx, w):
        (x,) = ctx.saved_tensors
        out = torch.empty_like(x)
        eng = ctx.eng
        eng._bind()
        eng.ctx.vjp_device(x.data_ptr(), w.contiguous().data_ptr(), out.data_ptr())
        return out, None


def velocity(x: torch.Tensor, eng: Engine) -> torch.Tensor:
    """u = U(x) for x of shape (pcount, 3), float64 on the GPU."""
    return _Velocity.apply(x, eng)


def _periodic_axes(eng: Engine):
    return {"periodic": (0, 1, 2), "openx": (1, 2), "openxy": (2,), "open": ()}[eng.boundary]


def unwrap(eng: Engine, x: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """Ghost unwrap of get_ghost_p (periodic.f90:36-55): g += L if x-g > L/2, -= L if < -L/2."""
    L = eng.box_size
    out = g.clone()
    for k in _periodic_axes(eng):
        d = x[:, k] - g[:, k]
        out[:, k] = g[:, k] + L * (d > L / 2).to(g.dtype) - L * (d < -L / 2).to(g.dtype)
    return out


def wrap(eng: Engine, x: torch.Tensor) -> torch.Tensor:
    """enforce_periodic's wrap (periodic.f90:188-213); derivative is the identity."""
    L = eng.box_size
    out = x.clone()
    for k in _periodic_axes(eng):
        v = x[:, k]
        out[:, k] = v - L * (v > L / 2).to(x.dtype) + L * (v < -L / 2).to(x.dtype)
    return out


def ab_step(x, u, hist, dt: float, k: int):
    """The ladder of pmotion (timestep.f90:51-61) for step k = 0, 1, 2, ...
    (history zero at the start: Euler, AB2, AB3, then AB4)."""
    u1, u2, u3 = hist
    if k == 0:
        x = x + dt * u
    elif k == 1:
        x = (x + (1.5 * dt) * u) - (0.5 * dt) * u1
    elif k == 2:
        x = ((x + ((23. / 12.) * dt) * u) - ((4. / 3.) * dt) * u1) + ((5. / 12.) * dt) * u2
    else:
        x = x + (dt * (((55. * u - 59. * u1) + 37. * u2) - 9. * u3)) / 24.
    return x, (u, u1, u2)


def rollout(x0: torch.Tensor, eng: Engine, nsteps: int, dt: float,
            checkpoint_every: int = 20) -> torch.Tensor:
    """Positions after nsteps of velocity + AB ladder + wrap, checkpointed."""
    z = torch.zeros_like(x0)
    state = (x0, z, z, z)
    k = 0
    while k < nsteps:
        m = min(checkpoint_every, nsteps - k)

        def segment(x, u1, u2, u3, k0=k, m=m):
            hist = (u1, u2, u3)
            for q in range(m):
                u = velocity(x, eng)
                x, hist = ab_step(x, u, hist, dt, k0 + q)
                x = wrap(eng, x)
            return (x,) + tuple(hist)

        if torch.is_grad_enabled():
            state = checkpoint(segment, *state, use_reentrant=False)
        else:
            state = segment(*state)
        k += m
    return state[0]


def line_length(eng: Engine, x: torch.Tensor) -> torch.Tensor:
    """Total length sum_i |ghosti - x_i| (line.f90:43)."""
    g = unwrap(eng, x, x[eng.infront.clamp(min=0)])
    d = torch.linalg.vector_norm(g - x, dim=1)
    return torch.where(eng.active, d, torch.zeros_like(d)).sum()


def energy(eng: Engine, x: torch.Tensor) -> torch.Tensor:
    """energy_info (diagnostic.f90:259-278) with u = U(x)."""
    u = velocity(x, eng)
    inf = eng.infront.clamp(min=0)
    beh = eng.behind.clamp(min=0)
    gi = unwrap(eng, x, x[inf])
    gb = unwrap(eng, x, x[beh])
    di = torch.linalg.vector_norm(x - gi, dim=1, keepdim=True)
    db = torch.linalg.vector_norm(x - gb, dim=1, keepdim=True)
    sdot = ((db * gi + (di - db) * x) - di * gb) / ((2. * db) * di)   # get_deriv_1
    term = (u * torch.cross(x, sdot, dim=1)).sum(dim=1)
    e = 0.5 * di[:, 0] * (term + term[inf])
    return eng.quant_circ * torch.where(eng.active, e, torch.zeros_like(e)).sum()


def gradient(objective: str, x0, eng: Engine, nsteps: int, dt: float,
             checkpoint_every: int = 20):
    """(J, dJ/dx0) for J = line length or energy of the positions after nsteps."""
    x = torch.as_tensor(x0, dtype=torch.float64, device="cuda").clone().requires_grad_(True)
    xf = rollout(x, eng, nsteps, dt, checkpoint_every)
    J = line_length(eng, xf) if objective == "length" else energy(eng, xf)
    (g,) = torch.autograd.grad(J, x)
    return float(J.detach()), g.detach().cpu().numpy()


def objective_value(objective: str, x0, eng: Engine, nsteps: int, dt: float) -> float:
    with torch.no_grad():
        x = torch.as_tensor(x0, dtype=torch.float64, device="cuda")
        xf = rollout(x, eng, nsteps, dt)
        J = line_length(eng, xf) if objective == "length" else energy(eng, xf)
    return float(J)


def dt_max(delta: float, quant_circ: float = 9.97e-4, corea: float = 1.3e-8) -> float:
    """The stability bound of timestep_check (initial.f90:606-622)."""
    dmin = delta / 2.
    return abs(dmin ** 2 / (quant_circ * math.log(dmin / (math.pi * corea))))
