"""The reference's own calls, run unchanged against this package on the GPU.

* tests/test_core.py:8-41 of the reference: GHZ with the default (numpy)
  backend, read with numpy; ⟨ZZ⟩ and ⟨Z⟩ without an explicit run().
* examples/state_prep_with_ad.ipynb (cells 1-2): the infidelity of
  two_local_circ(skeleton, params) against a target, backend="torch",
  cost.backward().  The gradient is checked against the oracle's fourth-order
  central differences (complex128 numpy, kak centre by scipy expm) and against
  the restated torch tape (qaravan_amd.torch_tape.tape_run) at n = 4, 12 and
  20.  The notebook's skeleton call brickwall_skeleton(n, num_layers) is, with
  the reference's current signature (core/skeletons.py:49), a 2D n x layers
  lattice; the n-qubit chain with two brickwork layers it means is
  brickwall_skeleton(n, 1, 2), used here.  torch.dot refuses mixed complex
  dtypes in this torch, so the target is cast to the state's dtype.
* applications/compilation.py:73-90 (msp_infidelity): U gates with torch
  angles, CNOTs and an initial state array.

Tolerances: relative gradient error 1e-10 for complex128 and 1e-5 for
complex64 (the state tolerance); the fourth-order differences carry about
1e-12 of truncation and rounding error at h = 1e-3.
"""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd import CNOT, Circuit, H, StatevectorSim, two_local_circ
from qaravan_amd.param_gates import U
from qaravan_amd.skeletons import brickwall_skeleton
from qaravan_amd.torch_tape import tape_run

pytestmark = pytest.mark.gpu
C64, C128 = torch.complex64, torch.complex128


def _ghz():
    return Circuit([H(0), CNOT([1, 0], 3), CNOT([2, 1], 3)], 3)


def test_reference_core_statevector_sim(gpu):
    # tests/test_core.py:8-24, unchanged but for the import
    sim = StatevectorSim(_ghz())
    sim.run()
    statevector = sim.get_statevector()
    expected = np.zeros(8, dtype=complex)
    expected[0] = 1 / np.sqrt(2)
    expected[-1] = 1 / np.sqrt(2)
    assert isinstance(statevector, np.ndarray) and statevector.dtype == np.complex128
    assert np.allclose(statevector, expected, atol=1e-7), "Statevector is not the expected GHZ state"
    assert isinstance(sim.state, np.ndarray) and sim.state.shape == (2, 2, 2)


def test_reference_core_pauli_expectation(gpu):
    # tests/test_core.py:26-41
    sim = StatevectorSim(_ghz())
    izz_exp = sim.pauli_expectation('izz')
    assert abs(izz_exp - 1.0) < 1e-7, f"Expected <Z_1 Z_2> = 1, got {izz_exp}"
    izi_exp = sim.pauli_expectation('izi')
    assert abs(izi_exp - 0.0) < 1e-7, f"Expected <Z_1> = 0, got {izi_exp}"


def test_backends_return_kinds(gpu):
    ref = O.run(_ghz().gate_list, 3)
    out = {}
    for backend, dev in (("numpy", "cpu"), ("torch", "cpu"), ("torch", "cuda"), ("hip", "cuda")):
        sim = StatevectorSim(_ghz(), backend=backend, device=dev)
        st = sim.run(progress_bar=False)
        sv = sim.get_statevector()
        if backend == "numpy":
            assert isinstance(st, np.ndarray) and st.shape == (2, 2, 2) and sv.shape == (8,)
            got = sv
        else:
            assert torch.is_tensor(sv) and sv.device.type == torch.device(dev).type
            assert tuple(st.shape) == (2, 2, 2)
            assert sv.dtype == C64
            got = sv.detach().cpu().numpy()
        out[(backend, dev)] = got
        assert np.linalg.norm(got - ref) < (1e-12 if backend == "numpy" else 1e-6)


def test_hip_backend_refuses_grad_matrices(gpu):
    p = torch.randn(15, dtype=torch.float64, requires_grad=True)
    circ = two_local_circ([(0, 1)], p)
    with pytest.raises(NotImplementedError):
        StatevectorSim(circ, backend="hip", device=gpu).run(progress_bar=False)


def _infidelity_torch(params, skeleton, target, dtype):
    # examples/state_prep_with_ad.ipynb, cell 1
    circ = two_local_circ(skeleton, params)
    sim = StatevectorSim(circ, backend="torch", dtype=dtype)
    sim.run(progress_bar=False)
    ansatz_sv = sim.get_statevector()
    return 1 - torch.abs(torch.dot(ansatz_sv.conj(), target.to(ansatz_sv.dtype))) ** 2


@pytest.mark.parametrize("n", [4, 12, 20])
@pytest.mark.parametrize("dtype", [C128, C64])
def test_infidelity_gradient_vs_fd_and_tape(gpu, n, dtype):
    rng = np.random.default_rng(100 + n)
    skeleton = brickwall_skeleton(n, 1, 2)
    npar = 15 * len(skeleton)
    p0 = rng.normal(size=npar)
    t = rng.random(2 ** n) + 1j * rng.random(2 ** n)
    t /= np.linalg.norm(t)
    target = torch.tensor(t)
    print(f"infidelity gradient n={n} {dtype} params={npar}", flush=True)
    params = torch.nn.Parameter(torch.tensor(p0))
    cost = _infidelity_torch(params, skeleton, target, dtype)
    cost.backward()
    g = params.grad.numpy()
    c_ref = O.infidelity(p0, skeleton, t, n)
    assert abs(cost.item() - c_ref) < (1e-12 if dtype == C128 else 1e-5)
    tol = 1e-10 if dtype == C128 else 1e-5
    # oracle: fourth-order central differences on a seeded subset of parameters
    idx = sorted(rng.choice(npar, size=min(npar, 12 if n < 20 else 6), replace=False).tolist())
    g_fd = -O.fd4_gradient(lambda x: O.fidelity(x, skeleton, t, n), p0, idx)
    err_fd = np.linalg.norm(g[idx] - g_fd) / np.linalg.norm(g_fd)
    # the restated torch tape (complex128 on the GPU), every parameter
    pt = torch.tensor(p0, requires_grad=True)
    circ = two_local_circ(skeleton, pt).build()
    sv = tape_run([gg for layer in circ.layers for gg in layer], n, gpu, C128)
    c_t = 1 - torch.abs(torch.dot(sv.conj(), target.to(gpu))) ** 2
    c_t.backward()
    g_tape = pt.grad.numpy()
    err_tape = np.linalg.norm(g - g_tape) / np.linalg.norm(g_tape)
    print(f"  rel err vs fd4 {err_fd:.2e}, vs tape {err_tape:.2e}", flush=True)
    assert err_fd < tol and err_tape < tol


def test_state_and_initial_state_gradients(gpu):
    # d/dψ_0 of a real loss through a fixed circuit: the adjoint of the run
    n = 6
    rng = np.random.default_rng(3)
    skeleton = brickwall_skeleton(n, 1, 2)
    circ = two_local_circ(skeleton, torch.tensor(rng.normal(size=15 * len(skeleton))))
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi0 = torch.tensor(v / np.linalg.norm(v), requires_grad=True)
    w = torch.tensor(rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n))
    sim = StatevectorSim(circ, init_state=psi0, backend="torch", dtype=C128)
    sim.run(progress_bar=False)
    loss = torch.abs(torch.dot(sim.get_statevector().conj(), w)) ** 2
    loss.backward()
    # tape restatement
    p2 = psi0.detach().clone().requires_grad_(True)
    sv = tape_run([gg for layer in circ.build().layers for gg in layer], n, gpu, C128, init=p2)
    l2 = torch.abs(torch.dot(sv.conj(), w.to(gpu))) ** 2
    l2.backward()
    assert abs(loss.item() - l2.item()) < 1e-12
    assert torch.linalg.vector_norm(psi0.grad - p2.grad).item() < 1e-11 * torch.linalg.vector_norm(p2.grad).item()


def test_msp_infidelity_u_gates_and_init_state(gpu):
    # applications/compilation.py:73-90: U gates with torch angles, CNOTs, an
    # initial state array, backend="torch"
    n = 5
    rng = np.random.default_rng(11)
    skeleton = brickwall_skeleton(n, 1, 1)
    npar = 6 * len(skeleton) + 3 * n
    p0 = rng.normal(size=npar)
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    in_state = v / np.linalg.norm(v)
    w = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    out_state = w / np.linalg.norm(w)

    def gates(p):
        gl, c = [], 0
        for pair in skeleton:
            gl.append(U(pair[0], *p[3 * c:3 * c + 3]))
            gl.append(U(pair[1], *p[3 * c + 3:3 * c + 6]))
            c += 2
            gl.append(CNOT(pair, n))
        for i in range(n):
            gl.append(U(i, *p[3 * c:3 * c + 3]))
            c += 1
        return gl

    params = torch.nn.Parameter(torch.tensor(p0))
    sim = StatevectorSim(Circuit(gates(params), n), backend="torch", init_state=in_state, dtype=C128)
    sim.run(progress_bar=False)
    sv = sim.get_statevector()
    cost = 1 - torch.abs(torch.dot(sv.conj(), torch.tensor(out_state))) ** 2
    cost.backward()

    def fid(p):
        psi = O.run(gates(p), n, in_state)
        return abs(np.vdot(psi, out_state)) ** 2
    assert abs(cost.item() - (1 - fid(p0))) < 1e-12
    idx = list(range(0, npar, 3))
    g_fd = -O.fd4_gradient(fid, p0, idx)
    err = np.linalg.norm(params.grad.numpy()[idx] - g_fd) / np.linalg.norm(g_fd)
    assert err < 1e-10


def test_op_action_kinds_and_grad(gpu):
    from qaravan_amd.statevector_sim import op_action
    rng = np.random.default_rng(2)
    n = 5
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    m = O.kak_unitary(rng.normal(size=15))
    ref = O.op_action(m, [3, 1], v)
    got = op_action(m, [3, 1], v)
    assert isinstance(got, np.ndarray) and np.abs(got - ref).max() < 1e-13
    got_r = op_action(m, [3, 1], v.reshape([2] * n))
    assert got_r.shape == (2,) * n and np.abs(got_r.reshape(-1) - ref).max() < 1e-13
    mt = torch.tensor(m, requires_grad=True)
    out = op_action(mt, [3, 1], torch.tensor(v))
    assert out.device.type == "cpu"
    out.abs().pow(2).mul(torch.arange(2 ** n)).sum().backward()
    # tape restatement of the same product
    mt2 = torch.tensor(m, requires_grad=True)
    from qaravan_amd.torch_tape import op_action as tape_op
    o2 = tape_op(mt2, [3, 1], torch.tensor(v).reshape([2] * n)).reshape(-1)
    o2.abs().pow(2).mul(torch.arange(2 ** n)).sum().backward()
    assert torch.abs(mt.grad - mt2.grad).max().item() < 1e-11


def test_non_unitary_gate_refuses_backward(gpu):
    from qaravan_amd import Gate
    p = torch.randn(15, dtype=torch.float64, requires_grad=True)
    circ = Circuit(two_local_circ([(0, 1)], p).gate_list + [Gate("P0", [1], np.array([[1, 0], [0, 0]]))], 2)
    sim = StatevectorSim(circ, backend="torch", dtype=C128)
    sim.run(progress_bar=False)
    loss = sim.get_statevector().abs().pow(2).sum()
    with pytest.raises(NotImplementedError):
        loss.backward()
