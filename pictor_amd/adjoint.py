"""Adjoint of the 2-D particle step exposed to PyTorch.

`ParticleStepObjective` is a torch.autograd.Function whose forward runs
nsteps steps of the particle step from the given initial momenta (and fixed
positions) and returns the objective L = 1/2 sum B^2 over the physical cells
at the final step; its backward is the hand-written adjoint in
libpictor_hip.so (segment checkpointing, recomputation of the particle state
inside each segment).  Both run through the C ABI on the GPU; torch tensors
are only the interface.

`drift_boost` is AddDriftVel (prob_dist.F90:223-254) for a drift along x
written with torch operations, so that a gradient with respect to the drift
speed follows from the momentum gradient by automatic differentiation:
    u = gamma_d (u_th + s beta_d gamma_th),  v = v_th,  w = w_th,
gamma_th = sqrt(1 + |u_th|^2), gamma_d = 1 / sqrt(1 - beta_d^2), s = +-1.
"""
import numpy as np
import torch

from .simulation import Simulation, setup_weibel


def drift_signs(n, epc):
    """Drift direction (+1 / -1) of every particle of setup_weibel's order:
    particle 2p + f (flavour f + 1) of position p; call = (p // (epc/2)) % 2;
    call 0 drifts flavour 1 by -drift and flavour 2 by +drift, call 1 the
    opposite."""
    idx = np.arange(n)
    p, f = idx // 2, idx % 2
    call = (p // (epc // 2)) % 2
    return np.where((call == 0) == (f == 0), -1.0, 1.0)


def drift_boost(u_th, v_th, w_th, sign, drift):
    """Boosted momenta (u, v, w); v and w are copies of v_th, w_th, so that a
    gradient retained on them is the derivative at fixed u (u depends on
    v_th and w_th through gamma_th)."""
    gth = torch.sqrt(1.0 + u_th * u_th + v_th * v_th + w_th * w_th)
    gd = 1.0 / torch.sqrt((1.0 - drift) * (1.0 + drift))
    u = gd * (u_th + sign * drift * gth)
    return u, v_th.clone(), w_th.clone()


class ParticleStepObjective(torch.autograd.Function):
    """L = 1/2 sum B^2 after `nsteps` steps of `sim` (a 2-D Simulation) from the
    particles `P` (positions, charges, flavours) with initial momenta u, v, w."""

    @staticmethod
    def forward(ctx, u, v, w, sim, P, nsteps, segment):
        ctx.sim = sim
        ctx.dev = u.device
        ctx.on_device = u.device.type == "cuda"
        if ctx.on_device:
            # momenta and particles stay on the GPU: the library reads them there
            dev = [t.detach().to(torch.float64).contiguous() for t in (u, v, w)]
            pd = {k: torch.as_tensor(P[k], device=u.device).to(torch.float64 if k != "flv" else torch.int32)
                  .contiguous() for k in ("q", "x", "y", "z", "flv")}
            torch.cuda.synchronize(u.device)
            L = sim.adjoint_forward_device(dev[0].numel(), pd["q"].data_ptr(), pd["x"].data_ptr(),
                                           pd["y"].data_ptr(), pd["z"].data_ptr(), dev[0].data_ptr(),
                                           dev[1].data_ptr(), dev[2].data_ptr(), pd["flv"].data_ptr(), nsteps,
                                           segment)
            ctx.n = dev[0].numel()
        else:
            arr = [t.detach().to("cpu", torch.float64).numpy() for t in (u, v, w)]
            L = sim.adjoint_forward(P["q"], P["x"], P["y"], P["z"], arr[0], arr[1], arr[2], P["flv"], nsteps,
                                    segment)
        return torch.tensor(L, dtype=torch.float64, device=u.device)

    @staticmethod
    def backward(ctx, gL):
        if ctx.on_device:
            out = [torch.empty(ctx.n, dtype=torch.float64, device=ctx.dev) for _ in range(3)]
            torch.cuda.synchronize(ctx.dev)
            ctx.sim.adjoint_backward_device(out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), float(gL))
        else:
            du, dv, dw = ctx.sim.adjoint_backward(float(gL))
            out = [torch.from_numpy(a).to(ctx.dev) for a in (du, dv, dw)]
        return out[0], out[1], out[2], None, None, None, None


def weibel_gradients(nx, ny, epc, drift, uth, seed, nsteps, segment, form="fp32_split", device="cpu"):
    """Objective and its gradients with respect to the drift speed and all
    initial momenta for the 2-D Weibel load of setup_weibel."""
    P = setup_weibel(nx, ny, epc=epc, drift=0.0, uth=uth, seed=seed)
    sgn = torch.from_numpy(drift_signs(len(P["x"]), epc)).to(device)
    d = torch.tensor(float(drift), dtype=torch.float64, device=device, requires_grad=True)
    th = [torch.from_numpy(P[k]).to(device).requires_grad_(True) for k in ("u", "v", "w")]
    u, v, w = drift_boost(th[0], th[1], th[2], sgn, d)
    for t in (u, v, w):
        t.retain_grad()
    # on a GPU device the particle arrays go there once, as the momenta do
    Pd = P if torch.device(device).type == "cpu" else {k: torch.from_numpy(a).to(device) for k, a in P.items()}
    with Simulation(nx, ny, epc=epc, form=form) as sim:
        L = ParticleStepObjective.apply(u, v, w, sim, Pd, nsteps, segment)
        L.backward()
        stats = sim.adjoint_stats()
    return {"objective": float(L), "d_drift": float(d.grad), "du": u.grad, "dv": v.grad, "dw": w.grad,
            "stats": stats}
