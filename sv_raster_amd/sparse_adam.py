"""SparseAdam with the reference's constructor and step semantics
(backends/new_cuda/new_svraster_cuda/sparse_adam.py:13-72): one HIP Adam
update per parameter tensor; `sparse=True` leaves elements whose gradient is
exactly zero untouched (adam_step.cu:42-45)."""
import torch

from . import _C


class SparseAdam(torch.optim.Optimizer):

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-15, biased=False, sparse=False):
        if not 0.0 <= lr:
            raise ValueError("Invalid learning rate: {}".format(lr))
        if not 0.0 <= eps:
            raise ValueError("Invalid epsilon value: {}".format(eps))
        if not 0.0 <= betas[0] < 1.0:
            raise ValueError("Invalid beta parameter at index 0: {}".format(betas[0]))
        if not 0.0 <= betas[1] < 1.0:
            raise ValueError("Invalid beta parameter at index 1: {}".format(betas[1]))
        defaults = dict(lr=lr, betas=betas, eps=eps)
        super(SparseAdam, self).__init__(params, defaults)
        self.biased = biased
        self.sparse = sparse

    def __setstate__(self, state):
        super(SparseAdam, self).__setstate__(state)

    @torch.no_grad()
    def use_device_steps(self):
        """Switch to the capture-safe update (no reference counterpart): each parameter's step count moves to
        an int32 device tensor that the update itself increments, so a captured step can be replayed.  The
        optimizer state (moments, step count) is created here, outside any capture.  Unbiased form only."""
        if self.biased:
            raise NotImplementedError("device steps are implemented for the unbiased update")
        for group in self.param_groups:
            for param in group['params']:
                state = self.state[param]
                if len(state) == 0:
                    state['step'] = 0
                    state['exp_avg'] = torch.zeros_like(param, memory_format=torch.preserve_format)
                    state['exp_avg_sq'] = torch.zeros_like(param, memory_format=torch.preserve_format)
                if 'step_dev' not in state:
                    state['step_dev'] = torch.full((1,), int(state['step']), dtype=torch.int32, device=param.device)
                    state['coef_dev'] = torch.zeros(2, dtype=torch.float32, device=param.device)
        self.device_steps = True

    @torch.no_grad()
    def step(self):
        if getattr(self, 'device_steps', False):
            for group in self.param_groups:
                lr = group['lr']
                beta1, beta2 = group['betas']
                for param in group['params']:
                    if param.grad is None:
                        continue
                    state = self.state[param]
                    _C.unbiased_adam_step_device(self.sparse, param, param.grad, state['exp_avg'],
                                                 state['exp_avg_sq'], state['step_dev'], state['coef_dev'], lr,
                                                 beta1, beta2, group['eps'])
            return
        for group in self.param_groups:
            lr = group['lr']
            beta1, beta2 = group['betas']
            eps = group['eps']
            for param in group['params']:
                if param.grad is None:
                    continue
                state = self.state[param]
                if len(state) == 0:
                    state['step'] = 0
                    state['exp_avg'] = torch.zeros_like(param, memory_format=torch.preserve_format)
                    state['exp_avg_sq'] = torch.zeros_like(param, memory_format=torch.preserve_format)
                state['step'] += 1
                if self.biased:
                    _C.biased_adam_step(self.sparse, param, param.grad, state['exp_avg'], state['exp_avg_sq'],
                                        lr, beta1, beta2, eps)
                else:
                    _C.unbiased_adam_step(self.sparse, param, param.grad, state['exp_avg'], state['exp_avg_sq'],
                                          state['step'], lr, beta1, beta2, eps)
