"""Batched QAOA on weighted Ising models over the HIP engine.

`IsingInstance` holds one model (IsingModel.h:28-212 semantics: E(s) =
-(sum h z + sum J z z), z = +1 for bit 0). `QaoaBatch` evaluates many
instances of uneven size at once on the GPU (qcs_qaoa_*), each with the
circuit of QAOAIsingSubalgorithm::Exec (IsingModel.h:557-567) and its exact
FP64 energy. `plain_optimize` is the reference's optimiser,
GradientDescentStep + Execute (IsingModel.h:220-280, 580-634: central
differences with eps, ADAM-W with beta1 0.8, beta2 0.95, lambda 1e-3, step
5e-4, stop when |E - E_old| <= 5e-4), run on exact energies instead of
100000-shot estimates; it counts exact statevector evaluations.
"""
from __future__ import annotations

import ctypes as C
import itertools
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib

_PD = C.POINTER(C.c_double)
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)


@dataclass
class IsingInstance:
    n: int
    edges: list  # (i, j) with i < j
    J: np.ndarray
    h: np.ndarray
    meta: dict = field(default_factory=dict)

    def energy(self, s: int) -> float:
        z = [1 - 2 * ((s >> i) & 1) for i in range(self.n)]
        e = sum(self.h[i] * z[i] for i in range(self.n))
        e += sum(Jk * z[i] * z[j] for (i, j), Jk in zip(self.edges, self.J))
        return -e

    def energies(self) -> np.ndarray:
        """E(s) for every basis state (vectorised)."""
        s = np.arange(1 << self.n, dtype=np.int64)
        z = 1 - 2 * ((s[:, None] >> np.arange(self.n)) & 1)
        e = z @ self.h
        for (i, j), Jk in zip(self.edges, self.J):
            e = e + Jk * z[:, i] * z[:, j]
        return -e


def random_regular_graph(n: int, d: int, rng: np.random.Generator, tries: int = 1000):
    """Uniform-ish random d-regular simple graph by the pairing model with
    rejection (n * d must be even)."""
    assert (n * d) % 2 == 0
    for _ in range(tries):
        stubs = np.repeat(np.arange(n), d)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        if np.any(pairs[:, 0] == pairs[:, 1]):
            continue
        es = {(int(min(a, b)), int(max(a, b))) for a, b in pairs}
        if len(es) == n * d // 2:
            return sorted(es)
    raise RuntimeError("no simple regular graph found")


def random_ising(n: int, rng: np.random.Generator, j_sigma=1.0, h_sigma=0.5) -> IsingInstance:
    """3-regular graph, J_ij ~ N(0, 1), h_i ~ N(0, 0.5) (0.5 is the standard
    deviation)."""
    edges = random_regular_graph(n, 3, rng)
    J = rng.normal(0.0, j_sigma, len(edges))
    h = rng.normal(0.0, h_sigma, n)
    return IsingInstance(n, edges, J, h)


PLAN_KINDS = {0: "lds", 1: "tile", 2: "mixer"}


def qaoa_plan(n: int, p: int, tile_bits: int = 12, group: int = 6):
    """The passes over the state that one evaluation of an n-qubit instance
    makes (qcs_qaoa_plan; host only), in order, as (kind, width, mode, qlo,
    layer) with kind "lds" (the whole evaluation in LDS), "tile" (mode 0:
    cost then Rx of `layer` on qubits 0..tile_bits-1; 1: Rx only) or "mixer"
    (qubits qlo..qlo+width-1; mode 0: Rx; 1: Rx then the energy; 2: Rx of
    `layer`, cost of layer + 1, Rx of layer + 1). The first pass starts from
    |+...+>."""
    cnt = C.c_uint32()
    check(lib().qcs_qaoa_plan(n, tile_bits, group, p, None, 0, C.byref(cnt)))
    buf = (C.c_uint32 * (5 * cnt.value))()
    check(lib().qcs_qaoa_plan(n, tile_bits, group, p, buf, cnt.value, C.byref(cnt)))
    w = list(buf)
    return [(PLAN_KINDS[w[5 * k]], *w[5 * k + 1 : 5 * k + 5]) for k in range(cnt.value)]


class QaoaBatch:
    """Many instances evaluated together on the GPU (qcs_qaoa_*)."""

    def __init__(self, instances, p: int, tile_bits: int = 12, group: int = 6, use_graph: bool = True):
        self.instances = list(instances)
        self.p = p
        nq = np.array([I.n for I in self.instances], dtype=np.uint32)
        ne = np.array([len(I.edges) for I in self.instances], dtype=np.uint32)
        ei = np.array([e[0] for I in self.instances for e in I.edges], dtype=np.uint32)
        ej = np.array([e[1] for I in self.instances for e in I.edges], dtype=np.uint32)
        J = np.concatenate([np.asarray(I.J, dtype=np.float64) for I in self.instances])
        h = np.concatenate([np.asarray(I.h, dtype=np.float64) for I in self.instances])
        self._keep = (nq, ne, ei, ej, J, h)
        hdl = C.c_void_p()
        check(
            lib().qcs_qaoa_create(
                len(self.instances), nq.ctypes.data_as(_PU32), ne.ctypes.data_as(_PU32), ei.ctypes.data_as(_PU32),
                ej.ctypes.data_as(_PU32), J.ctypes.data_as(_PD), h.ctypes.data_as(_PD), p, C.byref(hdl)
            )
        )
        self._h = hdl
        if (tile_bits, group, use_graph) != (12, 6, True):
            self.configure(tile_bits, group, use_graph)

    def close(self):
        if getattr(self, "_h", None):
            lib().qcs_qaoa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, tile_bits: int, group: int, use_graph: bool = True):
        check(lib().qcs_qaoa_configure(self._h, tile_bits, group, int(use_graph)))

    def evaluate(self, gammas, betas) -> np.ndarray:
        g = np.ascontiguousarray(np.broadcast_to(gammas, (len(self.instances), self.p)), dtype=np.float64)
        b = np.ascontiguousarray(np.broadcast_to(betas, (len(self.instances), self.p)), dtype=np.float64)
        out = np.zeros(len(self.instances))
        check(lib().qcs_qaoa_evaluate(self._h, g.ctypes.data_as(_PD), b.ctypes.data_as(_PD), out.ctypes.data_as(_PD)))
        return out

    def last_ms(self) -> float:
        ms = C.c_float()
        check(lib().qcs_qaoa_last_ms(self._h, C.byref(ms)))
        return float(ms.value)

    def state(self, i: int) -> np.ndarray:
        out = np.zeros(2 << self.instances[i].n)
        check(lib().qcs_qaoa_get_state(self._h, i, out.ctypes.data_as(_PD)))
        return out.view(np.complex128)

    def info(self) -> dict:
        st, tb, gr = C.c_uint32(), C.c_uint32(), C.c_uint32()
        amps = C.c_uint64()
        check(lib().qcs_qaoa_info(self._h, C.byref(st), C.byref(amps), C.byref(tb), C.byref(gr)))
        return {"stages": st.value, "amplitudes": amps.value, "tile_bits": tb.value, "group": gr.value}

    def autotune(self, gammas, betas, tiles=(10, 11, 12), groups=(4, 5, 6), reps: int = 2) -> dict:
        """Time an evaluation for every (tile_bits, group) and keep the fastest."""
        timings = {}
        for tb, gr in itertools.product(tiles, groups):
            try:
                self.configure(tb, gr, True)
                self.evaluate(gammas, betas)
                ms = min(self._timed(gammas, betas) for _ in range(reps))
                timings[(tb, gr)] = ms
            except Exception as exc:  # a configuration the device rejects
                timings[(tb, gr)] = float("inf")
                last_error = str(exc)  # noqa: F841
        best = min(timings, key=timings.get)
        self.configure(best[0], best[1], True)
        return {"best": best, "ms": {f"{k[0]}/{k[1]}": v for k, v in timings.items()}}

    def _timed(self, gammas, betas) -> float:
        self.evaluate(gammas, betas)
        return self.last_ms()


def plain_optimize(batch_factory, inst: IsingInstance, gamma0, beta0, eps=1e-3, alpha=5e-4, beta1=0.8, beta2=0.95,
                   lam=1e-3, delta_e=5e-4, max_iter=100000):
    """The reference's QAOA optimiser (IsingModel.h:220-280, 580-634) on exact
    energies. The 4p central-difference evaluations of a step plus the
    evaluation at the new angles run as one batch of 4p + 1 copies of the
    instance. Returns (best energy, best angles, exact evaluations, history)."""
    p = len(gamma0)
    batch = batch_factory([inst] * (4 * p), p)
    single = batch_factory([inst], p)
    gamma = np.array(gamma0, dtype=float)
    beta = np.array(beta0, dtype=float)
    mg, sg, mb, sb = (np.zeros(p) for _ in range(4))
    E_old, E = math.inf, 0.0
    E_min, best = math.inf, (gamma.copy(), beta.copy())
    evals, step, hist = 0, 0, []
    for it in range(max_iter):
        if not abs(E - E_old) > delta_e:
            break
        E_old = E
        step += 1
        div1 = 1.0 / (1.0 - beta1**step)
        div2 = 1.0 / (1.0 - beta2**step)
        G = np.tile(gamma, (4 * p, 1))
        B = np.tile(beta, (4 * p, 1))
        for i in range(p):
            G[4 * i, i] -= eps
            G[4 * i + 1, i] += eps
            B[4 * i + 2, i] -= eps
            B[4 * i + 3, i] += eps
        Es = batch.evaluate(G, B)
        evals += 4 * p
        ng, nb = gamma.copy(), beta.copy()
        for i in range(p):
            grad = (Es[4 * i + 1] - Es[4 * i]) / (2 * eps)
            mg[i] = (beta1 * mg[i] - (1 - beta1) * grad) * div1
            sg[i] = (beta2 * sg[i] + (1 - beta2) * grad * grad) * div2
            ng[i] = gamma[i] + alpha * (mg[i] / math.sqrt(sg[i] + 1e-8) - lam * gamma[i])
            grad = (Es[4 * i + 3] - Es[4 * i + 2]) / (2 * eps)
            mb[i] = (beta1 * mb[i] - (1 - beta1) * grad) * div1
            sb[i] = (beta2 * sb[i] + (1 - beta2) * grad * grad) * div2
            nb[i] = beta[i] + alpha * (mb[i] / math.sqrt(sb[i] + 1e-8) - lam * beta[i])
        gamma, beta = ng, nb
        E = float(single.evaluate(gamma[None, :], beta[None, :])[0])
        evals += 1
        hist.append(E)
        if E < E_min:
            E_min, best = E, (gamma.copy(), beta.copy())
    batch.close()
    single.close()
    return E_min, best, evals, hist
