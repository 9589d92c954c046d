"""Child-process check of the torch.autograd binding (run by
tests/test_gpu_adjoint.py): torch is imported first, then the library."""
import json
import os
import sys

import torch  # noqa: F401  (first: the library binds to torch's HIP runtime)

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy as np  # noqa: E402

import bgk3d_hit_amd as bgk  # noqa: E402
from bgk3d_hit_amd.torch_autograd import kinetic_energy  # noqa: E402

shape, omega, T = (16, 16, 16), 1.9, 20
u0n = bgk.hit_velocity(*shape, 2.0, 0.02, 3)
lat = bgk.Lattice(*shape, omega=omega, storage="f64")
J_ref, g_ref, _ = lat.adjoint_ke(u0n, T, ckpt_every=5, levels=1)
u0 = torch.tensor(u0n, device="cuda", requires_grad=True)
J = kinetic_energy(u0, lat, T, ckpt_every=5, levels=1)
(2.0 * J).backward()
print(json.dumps({"J_rel_err": abs(J.item() - J_ref) / J_ref,
                  "grad_equal": bool(np.array_equal(u0.grad.cpu().numpy(), 2.0 * g_ref))}))
