"""torch.autograd.Function over the residual's tangent and adjoint, for the
test harness only (the product exposes fcfd_jvp / fcfd_vjp through the
C-ABI; torch is plumbing here).

    R = ResidualFunction.apply(W, ns)      # W: [nv][N] float64 on the GPU

- forward: R(W) by fcfd_residual on the interior state;
- forward-mode AD (torch.autograd.forward_ad): the exact tangent
  dR = (dR/dW) dW by fcfd_jvp (FULL mode);
- reverse mode (backward / torch.autograd.grad): Wbar = (dR/dW)^T Rbar by
  fcfd_vjp.
Ghost slots are handled inside: the context's [nv][ld] buffers are padded
copies of the [nv][N] tensors torch sees.
"""
import torch

from freecfd_amd import _abi as A


def _pad(ns, X):
    N = ns.info.n_cells
    Xf = torch.zeros_like(ns.W)
    Xf[:, :N] = X
    return Xf


class ResidualFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, ns):
        N = ns.info.n_cells
        Wf = _pad(ns, W.detach())
        R = ns.assemble_residual(W=Wf, R=torch.zeros_like(ns.W))
        ctx.ns = ns
        ctx.save_for_backward(Wf)
        ctx.save_for_forward(Wf)
        return R[:, :N].clone()

    @staticmethod
    def backward(ctx, Rbar):
        ns = ctx.ns
        (Wf,) = ctx.saved_tensors
        N = ns.info.n_cells
        Wbar = ns.vjp(_pad(ns, Rbar.contiguous()), W=Wf.clone())
        return Wbar[:, :N].clone(), None

    @staticmethod
    def jvp(ctx, dW, _ns):
        ns = ctx.ns
        (Wf,) = ctx.saved_tensors
        N = ns.info.n_cells
        dR = ns.jvp(_pad(ns, dW.contiguous()), mode=A.JAC_JVP_FULL, W=Wf.clone())
        return dR[:, :N].clone()


def residual(W, ns):
    return ResidualFunction.apply(W, ns)


def jvp(ns, W, dW):
    """(R(W), dR/dW dW) through forward-mode AD"""
    import torch.autograd.forward_ad as fwAD
    with fwAD.dual_level():
        R = residual(fwAD.make_dual(W, dW), ns)
        p = fwAD.unpack_dual(R)
        return p.primal.clone(), p.tangent.clone()


def vjp(ns, W, Rbar):
    """(dR/dW)^T Rbar through reverse-mode AD"""
    Wr = W.detach().clone().requires_grad_(True)
    (g,) = torch.autograd.grad(residual(Wr, ns), Wr, grad_outputs=Rbar)
    return g
