"""The generic tile kernels (one compiled kernel for every pass structure,
stages run from the pattern library or interpreted) against the oracle.

Forward runs of random circuits with every op kind (dense 1- and 2-qubit,
CNOT, CZ, a 3-qubit dense gate) exercise the interpreted stages; the
hardware-efficient ansatz exercises the library patterns (forward and
adjoint).  Tolerances: ‖ψ − ψ_ref‖₂ ≤ 1e-12 (complex128), ≤ 1e-5 (complex64);
relative gradient error ≤ 1e-10 / 1e-5."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd import program as P
from qaravan_amd.workloads import hea_circuit, hea_theta, tfim_terms
from test_gpu_adjoint import hea_spec
from test_gpu_parity import random_circuit, sim_state

pytestmark = pytest.mark.gpu
C64, C128 = torch.complex64, torch.complex128
TOL = {C128: 1e-12, C64: 1e-5}

GCFGS = [P.TileConfig(tb=6, low_bits=4, max_cost=48, rb=3, generic=True, io_lane_bits=4),
         P.TileConfig(tb=9, low_bits=4, max_cost=1e9, rb=3, generic=True, io_lane_bits=4),
         P.TileConfig(tb=7, low_bits=4, max_cost=96, rb=4, generic=True, io_lane_bits=4),
         P.TileConfig(tb=8, low_bits=4, max_cost=1e9, rb=3, generic=True, io_lane_bits=4, remap=True),
         P.TileConfig(tb=6, low_bits=4, max_cost=64, rb=4, generic=True, io_lane_bits=4, remap=True)]


@pytest.mark.parametrize("ci", range(len(GCFGS)))
@pytest.mark.parametrize("dtype", [C64, C128])
def test_generic_random_circuits_vs_oracle(gpu, ci, dtype):
    cfg = GCFGS[ci]
    rng = np.random.default_rng(2000 + ci)
    n = cfg.tb + cfg.rb + 2 + (ci % 2)
    circ = random_circuit(n, 120, rng, k3=not cfg.remap)
    print(f"generic random n={n} cfg={cfg} dtype={dtype}", flush=True)
    sim, sv = sim_state(circ, gpu, dtype, cfg)
    prog = sim.program()
    assert all(p.flags & P.PASSF_GENERIC for p in prog.passes if p.kind == P.PASS_TILE)
    ref = O.run(circ.gate_list, n)
    err = np.linalg.norm(sv - ref)
    print(f"  |psi - psi_ref| = {err:.2e}", flush=True)
    assert err <= TOL[dtype]


@pytest.mark.parametrize("n,layers,tb,rb,remap", [(10, 3, 6, 3, False), (12, 4, 9, 3, False), (13, 5, 8, 3, True),
                                                   (14, 6, 9, 3, True), (11, 3, 7, 4, False)])
@pytest.mark.parametrize("dtype", [C128, C64])
def test_generic_adjoint_vs_oracle(gpu, n, layers, tb, rb, remap, dtype):
    from qaravan_amd.adjoint import AdjointModel
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    cfg = P.TileConfig(tb=tb, low_bits=4, max_cost=1e9, rb=rb, generic=True, io_lane_bits=4, remap=remap)
    print(f"generic adjoint n={n} layers={layers} tb={tb} rb={rb} remap={remap} {dtype}", flush=True)
    model = AdjointModel(circ, tfim_terms(n), len(theta), device=gpu, dtype=dtype, cfg=cfg)
    from qaravan_amd.specialize import _decode
    used = [int(st['pattern']) for prog in (model.pp.fwd, model.pp.adj) for st in _decode(prog.blob)[2]]
    if rb == 3:
        assert sum(u >= 0 for u in used) > 0, "library patterns expected for the ansatz"
    th = torch.tensor(theta, device=gpu, requires_grad=True)
    e = model(th)
    e.backward()
    e_ref, g_ref = O.adjoint_gradient(hea_spec(n, layers), theta, n, O.tfi_terms(n, -1.0, -1.0))
    g = th.grad.cpu().numpy()
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"  |E-E_ref|={abs(e.item() - e_ref):.2e} rel grad err={rel:.2e}", flush=True)
    assert abs(e.item() - e_ref) < (1e-11 if dtype == C128 else 1e-5)
    assert rel < (1e-10 if dtype == C128 else 1e-5)
