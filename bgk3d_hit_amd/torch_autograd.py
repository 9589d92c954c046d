"""PyTorch-ROCm binding of the discrete adjoint.

``kinetic_energy(u0, lattice, nsteps)`` returns J = sum_x |u(x)|^2/2 after
``nsteps`` BGK steps from the rho = 1 equilibrium of the velocity field u0
(a float64 CUDA tensor of shape (3, lz, ly, lx)); ``J.backward()`` fills
``u0.grad`` with the hand-written discrete adjoint (libbgk3d_hit.so).  The
adjoint pass also produces J, so forward() runs it once and backward() only
scales the stored gradient.  Torch is plumbing here: the computation is the
HIP library's.

HIP runtime sharing: the torch wheel carries its own libamdhip64.so.7.
Import this module (or torch) before the first call into libbgk3d_hit.so so
that the library binds to the runtime torch already loaded; the reverse
order leaves two HIP runtimes in the process and torch finds no GPU.
"""
import sys

if "bgk3d_hit_amd._lib" in sys.modules and sys.modules["bgk3d_hit_amd._lib"]._lib is not None \
        and "torch" not in sys.modules:
    raise ImportError("import torch (or bgk3d_hit_amd.torch_autograd) before the first use of "
                      "libbgk3d_hit.so: the library must bind to torch's HIP runtime")

import torch  # noqa: E402

from .lattice import Lattice  # noqa: E402


class KineticEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, lattice, nsteps, ckpt_every, levels):
        if u0.dtype != torch.float64 or not u0.is_cuda:
            raise TypeError("u0 must be a float64 CUDA tensor")
        lx, ly, lz = lattice.shape
        if tuple(u0.shape) != (3, lz, ly, lx):
            raise ValueError(f"u0 must have shape (3, {lz}, {ly}, {lx})")
        u = u0.detach().contiguous()
        torch.cuda.current_stream().synchronize()  # u0 is complete before the library reads it
        if ctx.needs_input_grad[0]:
            grad = torch.empty_like(u)
            J, stats = lattice.adjoint_ke_device(u.data_ptr(), grad.data_ptr(), nsteps,
                                                 ckpt_every, levels)
            ctx.save_for_backward(grad)
            ctx.stats = stats
        else:
            J = lattice.forward_ke(u.cpu().numpy(), nsteps)
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None, None


def kinetic_energy(u0, lattice: Lattice, nsteps, ckpt_every=25, levels=1):
    return KineticEnergy.apply(u0, lattice, nsteps, ckpt_every, levels)
