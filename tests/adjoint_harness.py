"""Verification harness for the adjoint: a torch.autograd.Function over the C-ABI
(adscf_adjoint).  Test infrastructure only; the product exposes the gradients
through the C-ABI and does not depend on torch."""
import numpy as np
import torch


class ScftLoss(torch.autograd.Function):
    """L(w_A, w_B, box) for a fixed chain and target; backward returns the adjoint gradients."""

    @staticmethod
    def forward(ctx_, w, box, scft, chain, phi_target, lam, ckpt):
        box_np = box.detach().cpu().numpy().astype(np.float64)
        scft.set_box(*box_np)
        out = scft.adjoint(chain[0], chain[1], w.detach().cpu().numpy().astype(np.float64), phi_target,
                           lam=lam, checkpoint_every=ckpt)
        ctx_.save_for_backward(torch.from_numpy(out["grad_w"]), torch.from_numpy(out["grad_box"]))
        return torch.tensor(out["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx_, grad_out):
        gw, gb = ctx_.saved_tensors
        return grad_out * gw, grad_out * gb, None, None, None, None, None


def loss_and_grads(scft, chain, w, box, phi_target, lam=0.1, ckpt=0):
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(box, dtype=torch.float64, requires_grad=True)
    L = ScftLoss.apply(wt, bt, scft, chain, phi_target, lam, ckpt)
    L.backward()
    return float(L.detach()), wt.grad.numpy().copy(), bt.grad.numpy().copy()


def loss_only(scft, chain, w, box, phi_target, lam=0.1, ckpt=0):
    scft.set_box(*box)
    return scft.adjoint(chain[0], chain[1], w, phi_target, lam=lam, checkpoint_every=ckpt, gradients=False)["loss"]
