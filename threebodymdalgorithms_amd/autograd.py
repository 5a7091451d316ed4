"""torch.autograd bindings of the energy forward and its hand-written adjoint.

    E = at_energy(pos, nu, engine)      pos: (N, 3) float64 CUDA tensor, nu: 0-d float64 tensor
    E.backward()                        pos.grad = dE/dx (the reference's f), nu.grad = dE/dnu

When pos requires a gradient the forward runs one adjoint sweep, which gives
E and dE/dx together, and the backward only scales the saved dE/dx; otherwise
the forward runs the energy-only sweep.  E is linear in nu, so dE/dnu is the
energy at nu = 1: E / nu, or at nu = 0 an energy sweep at nu = 1.  The sweeps
run on the engine's GPU through the C ABI on torch's current stream; the
engine's algorithm, cutoff, domain and precision apply.

`at_forces(pos, nu, engine)` returns dE/dx itself with a gradient for nu only
(dE/dx is linear in nu, so d(dE/dx)/dnu = (dE/dx) / nu, or dE/dx at nu = 1
when nu = 0); it serves force matching fits of nu.
"""
import torch

from .engine import Engine


def _stream_handle(device):
    return torch.cuda.current_stream(device).cuda_stream


def _check_pos(pos):
    if not isinstance(pos, torch.Tensor) or pos.dtype != torch.float64 or not pos.is_cuda:
        raise TypeError("pos must be a float64 CUDA tensor of shape (N, 3)")
    if pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError("pos must have shape (N, 3)")


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, nu, engine):
        _check_pos(pos)
        pos_c = pos.detach().contiguous()
        nu_v = float(nu)
        E = torch.zeros((), dtype=torch.float64, device=pos.device)
        grad = None
        if ctx.needs_input_grad[0]:
            # one sweep gives E and dE/dx together; backward only scales it
            grad = torch.empty_like(pos_c)
            engine.energy_grad_dev(pos_c.shape[0], pos_c.data_ptr(), nu_v, E.data_ptr(), grad.data_ptr(),
                                   _stream_handle(pos.device))
        else:
            engine.energy_dev(pos_c.shape[0], pos_c.data_ptr(), nu_v, E.data_ptr(), _stream_handle(pos.device))
        ctx.engine = engine
        ctx.nu_v = nu_v
        ctx.has_grad = grad is not None
        ctx.save_for_backward(pos_c, E.detach().clone(), grad if grad is not None else E.new_zeros(()))
        return E

    @staticmethod
    def backward(ctx, gE):
        pos_c, E, grad = ctx.saved_tensors
        g_pos = None
        if ctx.needs_input_grad[0]:
            if not ctx.has_grad:
                grad = torch.empty_like(pos_c)
                E2 = torch.zeros((), dtype=torch.float64, device=pos_c.device)
                ctx.engine.energy_grad_dev(pos_c.shape[0], pos_c.data_ptr(), ctx.nu_v, E2.data_ptr(),
                                           grad.data_ptr(), _stream_handle(pos_c.device))
            g_pos = grad * gE
        g_nu = None
        if ctx.needs_input_grad[1]:
            # E is linear in nu: dE/dnu is the energy at nu = 1
            if ctx.nu_v != 0.0:
                g_nu = (E / ctx.nu_v) * gE
            else:
                E1 = torch.zeros((), dtype=torch.float64, device=pos_c.device)
                ctx.engine.energy_dev(pos_c.shape[0], pos_c.data_ptr(), 1.0, E1.data_ptr(),
                                      _stream_handle(pos_c.device))
                g_nu = E1 * gE
        return g_pos, g_nu, None


class _Forces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, nu, engine):
        _check_pos(pos)
        pos_c = pos.detach().contiguous()
        nu_v = float(nu)
        E = torch.zeros((), dtype=torch.float64, device=pos.device)
        grad = torch.empty_like(pos_c)
        engine.energy_grad_dev(pos_c.shape[0], pos_c.data_ptr(), nu_v, E.data_ptr(), grad.data_ptr(),
                               _stream_handle(pos.device))
        ctx.nu_v = nu_v
        ctx.engine = engine
        ctx.save_for_backward(grad, pos_c)
        return grad

    @staticmethod
    def backward(ctx, gF):
        grad, pos_c = ctx.saved_tensors
        if ctx.needs_input_grad[0]:
            raise RuntimeError("at_forces has no second derivative with respect to positions")
        g_nu = None
        if ctx.needs_input_grad[1]:
            # dE/dx is linear in nu: its nu-derivative is dE/dx at nu = 1
            if ctx.nu_v != 0.0:
                g_nu = (gF * grad).sum() / ctx.nu_v
            else:
                E1 = torch.zeros((), dtype=torch.float64, device=pos_c.device)
                grad1 = torch.empty_like(pos_c)
                ctx.engine.energy_grad_dev(pos_c.shape[0], pos_c.data_ptr(), 1.0, E1.data_ptr(), grad1.data_ptr(),
                                           _stream_handle(pos_c.device))
                g_nu = (gF * grad1).sum()
        return None, g_nu, None


def at_energy(pos, nu, engine: Engine):
    """Total Axilrod-Teller energy of `pos` (differentiable in pos and nu)."""
    if not isinstance(nu, torch.Tensor):
        nu = torch.tensor(float(nu), dtype=torch.float64, device=pos.device)
    return _Energy.apply(pos, nu, engine)


def at_forces(pos, nu, engine: Engine):
    """dE/dx at `pos` (the reference's f), differentiable in nu."""
    if not isinstance(nu, torch.Tensor):
        nu = torch.tensor(float(nu), dtype=torch.float64, device=pos.device)
    return _Forces.apply(pos, nu, engine)
