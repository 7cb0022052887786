"""Environment-tensor circuit optimisation on the GPU.

Mirrors src/qaravan/tensorQ/environments.py:7-80:

* ``sv_environment(circ, left_sv, gate_idx)`` (:7-19): the 2^k x 2^k
  environment of gate ``gate_idx`` = partial_overlap(U_<g|0⟩, U_>g†|left⟩)
  over the gate's qubits, plus the gate's matrix.
* ``environment_update(circ, gate_idx, pre_state, post_state, target_sv,
  init_sv, direction)`` (:21-45): the gate that maximises |Tr(U·env)| is
  V·W† from env = W·S·V† (numpy SVD of the small environment on the host);
  the states move one gate along the sweep.  Returns (1 − |Tr(U·env)|,
  pre_state, post_state).
* ``environment_state_prep(target_sv, init_sv=None, circ=None, skeleton=None,
  context=None, quiet=True)`` (:47-80): alternating left/right sweeps.

The state vectors live on the GPU (complex128 by default: the optimisation
drives costs toward zero, where complex64 rounding would dominate); every
contraction and gate application is a library call (svh_partial_overlap,
svh_apply_dense).  Only the SVD of the 2^k x 2^k environment (4x4 for
2-qubit gates) runs on the host, as in the reference.  ``context`` is any
object with the fields the sweep loop reads (``resume``, ``run_state``,
``step``, ``max_iter``, ``step_update``); ``SweepContext`` is a minimal one
without checkpoint files.
"""
import numpy as np
import torch

from . import _lib
from .circuits import two_local_circ
from .measure import partial_overlap
from .statevector_sim import StatevectorSim, apply_matrix_inplace


class SweepContext:
    """Iteration bookkeeping of core/utils.py:144-232 (RunContext) without the
    checkpoint file: max_iter, optional plateau / absolute stop."""

    def __init__(self, max_iter=1000, convergence_check=False, stop_ratio=1e-6, stop_absolute=None,
                 resume=False, run_state=None):
        self.max_iter = max_iter
        self.convergence_check = convergence_check
        self.stop_ratio = stop_ratio
        self.stop_absolute = stop_absolute
        self.resume = resume
        self.run_state = run_state
        self.step = run_state.get('step', 0) if isinstance(run_state, dict) else 0
        self.term_msg = ""

    def step_update(self, run_state):
        self.run_state = run_state
        self.step = run_state.get('step', self.step)
        if self.step >= self.max_iter:
            self.term_msg = "Max iterations reached."
            return True
        if self.convergence_check:
            cl = run_state.get('cost_list')
            if cl is not None and len(cl) >= 2:
                if abs(cl[-1] - cl[-2]) < self.stop_ratio * abs(cl[-2]):
                    self.term_msg = f"Plateau detected with cost {cl[-1]} at step {self.step}"
                    return True
                if self.stop_absolute is not None and cl[-1] < self.stop_absolute:
                    self.term_msg = f"Absolute stop condition met with cost {cl[-1]} at step {self.step}"
                    return True
        return False


def _device_state(sv, device, dtype):
    t = sv if torch.is_tensor(sv) else torch.as_tensor(np.asarray(sv, dtype=np.complex128))
    return t.reshape(-1).to(device=device, dtype=dtype).clone()


def _polar_update(env):
    x, _, yh = np.linalg.svd(env)
    return yh.conj().T @ x.conj().T


def sv_environment(circ, left_sv, gate_idx, device="cuda", dtype=torch.complex128):
    """(environment, gate matrix) of gate ``gate_idx`` (environments.py:7-19)."""
    g = circ[gate_idx]
    sim1 = StatevectorSim(circ[:gate_idx], init_state=None, backend="hip", device=device, dtype=dtype)
    sim2 = StatevectorSim(circ[gate_idx + 1:].dag(), init_state=left_sv, backend="hip", device=device, dtype=dtype)
    sim1.run(progress_bar=False)
    sim2.run(progress_bar=False)
    env = partial_overlap(sim1.get_statevector(), sim2.get_statevector(), skip=list(g.indices))
    return env, g.matrix


def environment_update(circ, gate_idx, pre_state, post_state, target_sv, init_sv, direction='decreasing',
                       inplace=False):
    """One local update and one step of the sweep (environments.py:21-45).
    ``pre_state``/``post_state`` are flat device tensors; with ``inplace`` they
    are updated in place (the sweep driver owns them), otherwise copies are
    returned as the reference returns new arrays."""
    n = circ.num_sites
    indices = list(circ[gate_idx].indices)
    if not (torch.is_tensor(pre_state) and pre_state.is_cuda):
        pre_state = _device_state(pre_state, "cuda", torch.complex128)
    if not (torch.is_tensor(post_state) and post_state.is_cuda):
        post_state = _device_state(post_state, pre_state.device, pre_state.dtype)
    env = partial_overlap(pre_state, post_state, skip=indices).cpu().numpy()
    new_mat = _polar_update(env)
    circ.update_gate(gate_idx, new_mat)
    if not inplace:
        pre_state, post_state = pre_state.clone(), post_state.clone()
    if direction == 'decreasing':
        if gate_idx == 1:
            pre_state.copy_(torch.as_tensor(init_sv).reshape(-1))
        else:
            prev = circ[gate_idx - 1]
            apply_matrix_inplace(pre_state, n, np.asarray(prev.matrix).conj().T, list(prev.indices))
        apply_matrix_inplace(post_state, n, new_mat.conj().T, indices)
    else:
        apply_matrix_inplace(pre_state, n, new_mat, indices)
        if gate_idx == len(circ) - 2:
            post_state.copy_(torch.as_tensor(target_sv).reshape(-1))
        else:
            nxt = circ[gate_idx + 1]
            apply_matrix_inplace(post_state, n, np.asarray(nxt.matrix), list(nxt.indices))
    return 1 - np.abs(np.trace(new_mat @ env)), pre_state, post_state


def environment_state_prep(target_sv, init_sv=None, circ=None, skeleton=None, context=None, quiet=True,
                           device="cuda", dtype=torch.complex128):
    """Optimise ``circ`` (or a random 2-local circuit on ``skeleton``) so that
    circ|init⟩ ≈ |target⟩ by environment sweeps (environments.py:47-80).
    Returns (circ, cost_list)."""
    _lib.require_gpu()
    context = SweepContext() if context is None else context
    dev = torch.device(device)
    target = _device_state(target_sv, dev, dtype)
    if context.resume:
        circ = context.run_state['circ']
        cost_list = context.run_state['cost_list']
        pre_state = _device_state(context.run_state['pre_state'], dev, dtype)
        post_state = _device_state(context.run_state['post_state'], dev, dtype)
    else:
        if circ is None:
            circ = two_local_circ(skeleton)
        sim = StatevectorSim(circ, init_state=init_sv, backend="hip", device=dev, dtype=dtype)
        sim.run(progress_bar=False)
        ansatz = sim.get_statevector()
        from .statevector_sim import inner
        ov = complex(inner(target, ansatz, circ.num_sites, sim._scratch).item())
        cost_list = [1 - abs(ov)]
        pre_state = ansatz.clone()
        last = circ[-1]
        apply_matrix_inplace(pre_state, circ.num_sites, np.asarray(last.matrix).conj().T, list(last.indices))
        post_state = target.clone()
    n = circ.num_sites
    if init_sv is None:
        init = torch.zeros(2 ** n, dtype=dtype, device=dev)
        init[0] = 1.0
    else:
        init = _device_state(init_sv, dev, dtype)
    it = range(context.step, context.max_iter)
    if not quiet:
        from tqdm import tqdm
        it = tqdm(it)
    for step in it:
        for i in reversed(range(1, len(circ))):
            cost, pre_state, post_state = environment_update(circ, i, pre_state, post_state, target, init,
                                                             'decreasing', inplace=True)
            cost_list.append(cost)
        for i in range(len(circ) - 1):
            cost, pre_state, post_state = environment_update(circ, i, pre_state, post_state, target, init,
                                                             'increasing', inplace=True)
            cost_list.append(cost)
        run_state = {'step': step, 'circ': circ, 'cost_list': cost_list, 'pre_state': pre_state,
                     'post_state': post_state}
        if context.step_update(run_state):
            break
    return circ, cost_list
