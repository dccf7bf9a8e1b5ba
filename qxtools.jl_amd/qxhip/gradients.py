"""Gradients of amplitude losses with respect to gate angles.

`qx_amplitudes_vjp` returns, per data key, grad = sum_b amp_cot_b conj(d amp_b / dT)
(the ChainRulesCore pullback of a holomorphic map).  For a parameterised gate T(theta)
the chain rule gives dL/dtheta = Re(sum_i conj(grad_i) dT_i/dtheta) for a real loss L
whose cotangent is amp_cot = dL/d(Re amp) + i dL/d(Im amp), e.g. amp_cot = 2 amp for
L = sum |amp|^2.  dT/dtheta is the derivative matrix pushed through the same
circuit-to-tensor conversion as the gate itself (diagonal gates stay rank-1 on
their wire).  Every gate tensor gets its own data key (no dedupe).  No
counterpart exists in the reference, which is forward-only.
"""

from __future__ import annotations

import numpy as np

from . import _native
from .circuits import Circuit, gate_matrix, gate_matrix_derivative, gate_qubits
from .compute_graph import build_compute_graph
from .planning import contraction_scheme, sweep_contraction_plan
from .simulation import Executor
from .tn import TensorNetworkCircuit, gate_tensor_map


class ParamCircuit:
    """A parameterised circuit compiled to a saved plan with per-gate data keys."""

    def __init__(self, circ: Circuit, planner: str = "sweep", nslice: int = 0, precision: str = "c128",
                 batch: int | None = None, device: int = 0, graph: int = 1, **plan_kwargs):
        self.circ = circ
        tnc, names = gate_tensor_map(circ)
        tnc.add_input()
        tnc.add_output()
        if planner == "sweep":
            plan, groups = sweep_contraction_plan(tnc, **plan_kwargs), []
        else:
            groups, plan, _ = contraction_scheme(tnc, nslice, **plan_kwargs)
        self.graph = build_compute_graph(tnc, plan, groups, dedupe=False)
        key_of = {c.name: c.args[0] for c in self.graph.commands if c.kind == "load"}
        self._pgates = []   # (gate index, data key) of every parameterised gate
        for gi, (g, tn_names) in enumerate(zip(circ.ops, names)):
            if g.param is None:
                continue
            if len(tn_names) != 1:
                raise ValueError("parameterised gates must convert to one tensor")
            self._pgates.append((gi, key_of[tn_names[0]]))
        self._set_terms()
        self.dsl = self.graph.to_dsl()
        self.executor = Executor(self.dsl, self.graph.data, precision=precision, device=device, batch=batch,
                                 graph=graph)
        self.num_params = circ.num_params

    def _gate_tensor(self, g, deriv: bool) -> np.ndarray:
        """The gate's tensor (or dT/dtheta) through the same conversion as the network."""
        scratch = TensorNetworkCircuit(self.circ.num_qubits)
        m = gate_matrix_derivative(g) if deriv else gate_matrix(g)
        (name,) = scratch.push(gate_qubits(g), m)
        return scratch.tn[name].data

    def _set_terms(self) -> None:
        """Chain-rule terms (param index, key, dT/dtheta in the key's array layout) at
        the current angles, as one segmented reduction over all parameters."""
        self.terms = []
        for gi, key in self._pgates:
            g = self.circ.ops[gi]
            d = self._gate_tensor(g, True)
            if d.shape != self.graph.data[key].shape:
                raise ValueError(f"derivative of gate {g.name} has a different tensor structure")
            self.terms.append((g.param, key, d))
        self._tp = np.array([p for p, _, _ in self.terms], dtype=np.int64)
        self._tkeys = [k for _, k, _ in self.terms]
        self._td = np.concatenate([d.reshape(-1, order="F") for _, _, d in self.terms]) if self.terms else np.zeros(0)
        self._tstart = np.cumsum([0] + [d.size for _, _, d in self.terms[:-1]]).astype(np.int64)

    def params(self) -> np.ndarray:
        return self.circ.params()

    def set_params(self, theta) -> None:
        """Move the plan to new angles: every parameterised gate's tensor is rebuilt
        and re-uploaded into its device buffer (`qx_plan_update_tensor`), so the
        lowering, tuning and captured graphs are kept; the chain-rule terms follow."""
        theta = np.asarray(theta, dtype=np.float64)
        if theta.shape != (self.num_params,):
            raise ValueError(f"need {self.num_params} angles, got shape {theta.shape}")
        new = self.circ.with_params(theta)
        tensors = []
        for gi, key in self._pgates:
            t = TensorNetworkCircuit(new.num_qubits)
            (name,) = t.push(gate_qubits(new.ops[gi]), gate_matrix(new.ops[gi]))
            arr = t.tn[name].data
            if arr.shape != self.graph.data[key].shape:
                raise ValueError(f"gate {new.ops[gi].name} at the new angle has a different tensor structure")
            tensors.append((key, arr))
        for key, arr in tensors:
            self.executor.plan.update_tensor(key, arr)
            self.graph.data[key] = arr
        self.circ = new
        self._set_terms()

    def amplitudes(self, bits) -> np.ndarray:
        return self.executor.amplitudes(bits)

    def param_grad(self, key_grad: dict) -> np.ndarray:
        """dL/dtheta_p = sum over the gate tensors T of parameter p of
        Re <dL/dT, dT/dtheta_p> (the cotangent is conj-linear)."""
        out = np.zeros(self.num_params)
        if not self.terms:
            return out
        g = np.concatenate([key_grad[k] for k in self._tkeys])
        per_term = np.add.reduceat(np.real(np.conj(g) * self._td), self._tstart)
        np.add.at(out, self._tp, per_term)
        return out

    def _param_grad_flat(self, g_flat: np.ndarray) -> np.ndarray:
        """`param_grad` from the pullback's flat gradient (layout of the plan's
        `key_offsets`): the terms' entries gathered by one precomputed index."""
        out = np.zeros(self.num_params)
        if not self.terms:
            return out
        if getattr(self, "_tidx", None) is None:
            offs = self.executor.plan.key_offsets()
            self._tidx = np.concatenate([np.arange(offs[k][0], offs[k][0] + offs[k][1]) for k in self._tkeys])
        per_term = np.add.reduceat(np.real(np.conj(g_flat[self._tidx]) * self._td), self._tstart)
        np.add.at(out, self._tp, per_term)
        return out

    def loss_and_grad(self, bits) -> tuple:
        """L = sum_b |amp_b|^2 and dL/dtheta: one pass of forward and pullback when
        every slice variable is contracted inline, else the amplitudes first and the
        pullback of the cotangent 2 amp after."""
        try:
            amps, g_flat = self.executor.plan.norm_vjp(bits, flat=True)
        except _native.QxError as e:
            if e.status != _native.QX_ERR_UNSUPPORTED:
                raise
            amps = self.amplitudes(bits)
            kg = self.executor.vjp(bits, 2.0 * amps)
            return float(np.sum(np.abs(amps) ** 2)), self.param_grad(kg), amps
        loss = float(np.sum(np.abs(amps) ** 2))
        return loss, self._param_grad_flat(g_flat), amps
