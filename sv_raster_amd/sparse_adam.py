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

    # steps covered by each parameter's device table of bias terms (refilled ahead of use)
    BIAS_TABLE_STEPS = 1024

    @torch.no_grad()
    def use_device_steps(self, skip=None):
        """Switch to the capture-safe update (no reference counterpart): each parameter's step count moves to
        an int32 device tensor that the update itself increments, so a captured step can be replayed.  The
        bias terms of each step come from a device table filled on the host with the host-step formula, so the
        replayed update is bit-equal to the host-step one.  `skip` (an int32 [1] device tensor, optional): a
        nonzero value makes the update leave everything unchanged, the step count included.  The optimizer
        state (moments, step count) is created here, outside any capture.  Unbiased form only."""
        if self.biased:
            raise NotImplementedError("device steps are implemented for the unbiased update")
        for group in self.param_groups:
            beta1, beta2 = group['betas']
            for param in group['params']:
                state = self.state[param]
                if len(state) == 0:
                    state['step'] = 0
                    state['exp_avg'] = torch.zeros_like(param, memory_format=torch.preserve_format)
                    state['exp_avg_sq'] = torch.zeros_like(param, memory_format=torch.preserve_format)
                if 'step_dev' not in state:
                    state['step_dev'] = torch.full((1,), int(state['step']), dtype=torch.int32, device=param.device)
                    state['coef_dev'] = torch.zeros(3, dtype=torch.float32, device=param.device)
                    state['bias_table'] = _C.adam_bias_table(int(state['step']) + 1, self.BIAS_TABLE_STEPS,
                                                             group['lr'], beta1, beta2, param.device)
        self.device_skip = skip
        self._calls_since_fill = 0
        self.device_steps = True

    @torch.no_grad()
    def note_device_steps(self, n=1):
        """Records n device-step updates run outside step() (graph replays of a captured step).  When the
        tables of bias terms may run out, reads the device step counts (one synchronisation) and refills the
        tables in place from there, so replays keep taking host-form terms."""
        self._calls_since_fill += n
        if self._calls_since_fill < self.BIAS_TABLE_STEPS // 2:
            return
        for group in self.param_groups:
            beta1, beta2 = group['betas']
            for param in group['params']:
                state = self.state[param]
                if 'bias_table' not in state:
                    continue
                t = int(state['step_dev'].item())
                state['bias_table'].copy_(_C.adam_bias_table(t + 1, self.BIAS_TABLE_STEPS, group['lr'], beta1, beta2,
                                                             param.device))
        self._calls_since_fill = 0

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
                                                 beta1, beta2, group['eps'], bias_table=state['bias_table'],
                                                 skip_dev=self.device_skip)
            if not torch.cuda.is_current_stream_capturing():
                self.note_device_steps(1)
            return
        # every tensor's update (the reference's per-tensor step, adam_step.cu:106-162) in one launch per device
        items = {}
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
                items.setdefault(param.device, []).append(
                    (param, param.grad, state['exp_avg'], state['exp_avg_sq'], state['step'], lr, beta1, beta2, eps))
        for its in items.values():
            _C.adam_step_multi(self.biased, self.sparse, its)
