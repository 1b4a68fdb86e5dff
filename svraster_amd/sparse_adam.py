"""Adam optimizer over the voxel parameters, API-compatible with the reference's
cuda/svraster_cuda/sparse_adam.py:13-72.  Each parameter tensor is updated in place by one
launch of the library's Adam kernel (unbiased: bias corrections computed in double on the host;
biased: plain moments).  With sparse=True, elements whose gradient is exactly zero are left
untouched (moments included)."""
import torch

from . import _C


class SparseAdam(torch.optim.Optimizer):

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-15, biased=False, sparse=False):
        if not 0.0 <= lr:
            raise ValueError(f"Invalid learning rate: {lr}")
        if not 0.0 <= eps:
            raise ValueError(f"Invalid epsilon value: {eps}")
        if not 0.0 <= betas[0] < 1.0:
            raise ValueError(f"Invalid beta parameter at index 0: {betas[0]}")
        if not 0.0 <= betas[1] < 1.0:
            raise ValueError(f"Invalid beta parameter at index 1: {betas[1]}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps))
        self.biased = biased
        self.sparse = sparse

    @torch.no_grad()
    def step(self):
        for group in self.param_groups:
            lr = group["lr"]
            beta1, beta2 = group["betas"]
            eps = group["eps"]
            for param in group["params"]:
                if param.grad is None:
                    continue
                state = self.state[param]
                if len(state) == 0:
                    state["step"] = 0
                    state["exp_avg"] = torch.zeros_like(param, memory_format=torch.preserve_format)
                    state["exp_avg_sq"] = torch.zeros_like(param, memory_format=torch.preserve_format)
                state["step"] += 1
                if self.biased:
                    _C.biased_adam_step(self.sparse, param, param.grad, state["exp_avg"],
                                        state["exp_avg_sq"], lr, beta1, beta2, eps)
                else:
                    _C.unbiased_adam_step(self.sparse, param, param.grad, state["exp_avg"],
                                          state["exp_avg_sq"], state["step"], lr, beta1, beta2, eps)
