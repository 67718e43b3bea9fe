"""Learned angle proposals for QAOA on Ising instances.

An MLP (qcs_mlp_*: BF16 weights and activations, FP32 accumulation, trained
by the project's HIP forward/backward/Adam kernels) maps instance statistics
plus the 2p normalised angles to the energy, normalised by the instance's
scale S = sum|J| + sum|h|. It is trained on (instance, angles, exact energy)
tuples that the batched simulator (QaoaBatch) generates. For a new instance
it scores many random angle sets; the best `top` of them, in predicted
order, are re-scored exactly on the statevector a few at a time until a
stage brings no improvement (deployable_search), and the plain optimiser
(qaoa.plain_optimize) is the fallback when no proposal beats its default
starting point.

    python -m qcsim_amd.learned --train-rows 1000000 --sizes 16,18,20 --test-n 20
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time

import numpy as np

from ._lib import check, lib
from .qaoa import IsingInstance, QaoaBatch, plain_optimize, random_ising

_PF = C.POINTER(C.c_float)
N_STATS = 9


def instance_stats(inst: IsingInstance) -> np.ndarray:
    """Graph / J / h statistics (degree, coupling and field moments)."""
    J = np.asarray(inst.J, dtype=np.float64)
    h = np.asarray(inst.h, dtype=np.float64)
    deg = np.zeros(inst.n)
    for i, j in inst.edges:
        deg[i] += 1
        deg[j] += 1
    scale = np.sum(np.abs(J)) + np.sum(np.abs(h))
    return np.array(
        [inst.n / 24.0, deg.mean() / 3.0, J.mean(), J.std(), np.abs(J).mean(), h.mean(), h.std(), np.abs(h).mean(),
         scale / inst.n],
        dtype=np.float64,
    )


def energy_scale(inst: IsingInstance) -> float:
    return float(np.sum(np.abs(inst.J)) + np.sum(np.abs(inst.h)))


def features(stats: np.ndarray, gammas: np.ndarray, betas: np.ndarray) -> np.ndarray:
    """Rows of [stats, gamma/pi, beta/pi] (angles are taken in [0, pi))."""
    G = np.atleast_2d(gammas) / math.pi
    B = np.atleast_2d(betas) / math.pi
    S = np.broadcast_to(stats, (G.shape[0], stats.shape[0]))
    return np.concatenate([S, G, B], axis=1).astype(np.float32)


class Surrogate:
    """The MLP on the GPU (qcs_mlp_*)."""

    def __init__(self, in_dim: int, hidden: int = 256, layers: int = 3, seed: int = 1):
        h = C.c_void_p()
        check(lib().qcs_mlp_create(in_dim, hidden, layers, seed, C.byref(h)))
        self._h = h
        self.in_dim = in_dim

    def close(self):
        if getattr(self, "_h", None):
            lib().qcs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, X: np.ndarray, Y: np.ndarray, batch=4096, epochs=20, lr=1e-3, seed=3) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        check(lib().qcs_mlp_train(self._h, X.ctypes.data_as(_PF), Y.ctypes.data_as(_PF), X.shape[0], batch, epochs, lr,
                                  seed, losses.ctypes.data_as(_PF)))
        return losses

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        check(lib().qcs_mlp_forward(self._h, X.ctypes.data_as(_PF), X.shape[0], y.ctypes.data_as(_PF)))
        return y


def generate_dataset(sizes, n_instances: int, rows: int, p: int, rng: np.random.Generator, chunk_instances=256,
                     progress=None):
    """(features, normalised energies) from random instances and angles, all
    energies exact (QaoaBatch). Returns X, Y, seconds of GPU evaluation.
    `progress(text)` (optional) is called after each chunk of instances."""
    per_inst = max(1, rows // n_instances)
    Xs, Ys = [], []
    t_eval = 0.0
    for c0 in range(0, n_instances, chunk_instances):
        insts = [random_ising(int(rng.choice(sizes)), rng) for _ in range(min(chunk_instances, n_instances - c0))]
        stats = [instance_stats(I) for I in insts]
        scales = np.array([energy_scale(I) for I in insts])
        batch = QaoaBatch(insts, p)
        for _ in range(per_inst):
            G = rng.uniform(0, math.pi, (len(insts), p))
            B = rng.uniform(0, math.pi, (len(insts), p))
            t0 = time.perf_counter()
            E = batch.evaluate(G, B)
            t_eval += time.perf_counter() - t0
            Xs.append(np.concatenate([np.stack(stats), G / math.pi, B / math.pi], axis=1).astype(np.float32))
            Ys.append(E / scales)
        batch.close()
        if progress:
            progress(f"dataset: {c0 + len(insts)}/{n_instances} instances, {sum(len(y) for y in Ys)} rows, "
                     f"gpu eval {t_eval:.1f} s")
    X = np.concatenate(Xs, axis=0)
    Y = np.concatenate(Ys).astype(np.float32)
    return X, Y, t_eval


def propose(model: Surrogate, inst: IsingInstance, p: int, rng: np.random.Generator, candidates=65536, top=32):
    """Top-`top` angle sets by predicted energy among random candidates."""
    G = rng.uniform(0, math.pi, (candidates, p))
    B = rng.uniform(0, math.pi, (candidates, p))
    pred = model.predict(features(instance_stats(inst), G, B))
    idx = np.argsort(pred)[:top]
    return G[idx], B[idx], pred[idx] * energy_scale(inst)


def rescore(model: Surrogate, inst: IsingInstance, p: int, rng: np.random.Generator, top=32, candidates=65536):
    """The top proposals in predicted order with their exact energies
    (one batched evaluation of `top` copies of the instance)."""
    G, B, _ = propose(model, inst, p, rng, candidates, top)
    batch = QaoaBatch([inst] * top, p)
    E = batch.evaluate(G, B)
    batch.close()
    return G, B, E


def learned_search(model: Surrogate, inst: IsingInstance, p: int, rng: np.random.Generator, top=32, candidates=65536):
    """Propose, re-score all proposals exactly; returns (best E, angles, evals)."""
    G, B, E = rescore(model, inst, p, rng, top, candidates)
    k = int(np.argmin(E))
    return float(E[k]), (G[k], B[k]), top


def staged_rescore(inst: IsingInstance, p: int, G, B, batch=4, tol=1e-3, factory=None):
    """Re-score proposals in predicted order, `batch` at a time, and stop after
    the first batch (beyond the first) that lowers the running best by no
    more than `tol`. The rule sees only the proposals' own exact energies.
    Returns (energies re-scored, in order)."""
    factory = factory or (lambda insts, pp: QaoaBatch(insts, pp))
    ev = factory([inst] * batch, p)
    out = []
    best = math.inf
    for b0 in range(0, G.shape[0], batch):
        g, b = G[b0:b0 + batch], B[b0:b0 + batch]
        m = g.shape[0]
        if m < batch:  # a last, short batch: pad with copies of its first row
            g = np.concatenate([g, np.repeat(g[:1], batch - m, axis=0)])
            b = np.concatenate([b, np.repeat(b[:1], batch - m, axis=0)])
        E = ev.evaluate(g, b)[:m]
        out.extend(float(e) for e in E)
        new_best = min(best, float(E.min()))
        improved = best - new_best > tol
        best = new_best
        if not improved:
            break
    ev.close()
    return np.array(out)


def deployable_search(model: Surrogate, inst: IsingInstance, p: int, rng: np.random.Generator, top=32, batch=4,
                      tol=1e-3, candidates=65536, start=1.0, optimizer=None):
    """What a deployment spends: the top `top` proposals re-scored in stages
    (staged_rescore), one exact evaluation at the plain optimiser's default
    start (gamma = beta = `start`), and the plain optimiser from that start as
    the fallback when no proposal beats it. Nothing here looks at the plain
    optimiser's result. Returns a dict with the best energy, its angles and
    every exact evaluation counted."""
    factory = lambda insts, pp: QaoaBatch(insts, pp)  # noqa: E731
    G, B, _ = propose(model, inst, p, rng, candidates, top)
    E = staged_rescore(inst, p, G, B, batch, tol, factory)
    k = int(np.argmin(E))
    one = factory([inst], p)
    e_start = float(one.evaluate(np.full((1, p), start), np.full((1, p), start))[0])
    one.close()
    out = {"E": float(E[k]), "angles": (G[k], B[k]), "rescored": int(E.shape[0]), "evals": int(E.shape[0]) + 1,
           "fallback": False, "start_E": e_start, "proposals": (G, B)}
    if E[k] >= e_start:
        opt = optimizer or plain_optimize
        e_fb, ang, ev_fb, _ = opt(factory, inst, [start] * p, [start] * p)
        out["fallback"] = True
        out["evals"] += ev_fb
        if e_fb < out["E"]:
            out["E"], out["angles"] = e_fb, ang
    return out


def evals_to_target(energies, target: float, per_eval: int = 1):
    """Exact evaluations spent until the running best first reaches
    `target` (energies in evaluation order), or None."""
    best = math.inf
    for k, e in enumerate(energies):
        best = min(best, e)
        if best <= target:
            return (k + 1) * per_eval
    return None


# The strengthened plain baseline: the reference's loop (same gradient,
# ADAM-W moments, weight decay and stopping rule) with a step 40 times the
# reference's 5e-4, so that it leaves its starting point; at the reference
# step the loop stops after 7-9 iterations close to where it started.
LONG_STEP = 0.02


def long_optimize(factory, inst, gamma0, beta0):
    return plain_optimize(factory, inst, gamma0, beta0, alpha=LONG_STEP, max_iter=3000)


def compare_on_instance(model: Surrogate, inst: IsingInstance, p: int, rng: np.random.Generator, top=32,
                        gap=1e-3, batch=4, strengthened=True):
    """Plain optimiser against learned proposals on one instance.

    plain: the reference loop from gamma = beta = 1 on exact energies, 4p + 1
    evaluations per iteration. learned (deployable_search): proposals
    re-scored in stages with the fallback rule; its count is every exact
    evaluation it makes. Also reported: all `top` proposals re-scored
    (config's literal count), the evaluations until the first proposal
    reaches plain's best + gap (a count that needs plain's result, so no
    deployment could stop on it), and with `strengthened` the larger-step
    plain loop from the default start and warm-started from the learned
    best."""
    factory = lambda insts, pp: QaoaBatch(insts, pp)  # noqa: E731
    e_plain, _, ev_plain, hist = plain_optimize(factory, inst, [1.0] * p, [1.0] * p)
    plain_to_gap = evals_to_target(hist, e_plain + gap, 4 * p + 1)
    dep = deployable_search(model, inst, p, rng, top=top, batch=batch, tol=gap)
    # the same proposals, all re-scored at once
    G, B = dep["proposals"]
    ev = factory([inst] * G.shape[0], p)
    E_all = ev.evaluate(G, B)
    ev.close()
    out = {
        "n": inst.n,
        "ground_E": float(inst.energies().min()),
        "plain_E": e_plain,
        "plain_evals": ev_plain,
        "plain_evals_to_gap": plain_to_gap,
        "learned_E": dep["E"],
        "learned_evals": dep["evals"],
        "learned_rescored": dep["rescored"],
        "learned_fallback": dep["fallback"],
        "within_gap": bool(dep["E"] <= e_plain + gap),
        "eval_ratio": ev_plain / dep["evals"],
        "all_top_E": float(E_all.min()),
        "all_top_evals": top,
        "eval_ratio_all_top": ev_plain / top,
        "learned_evals_to_gap": evals_to_target(E_all, e_plain + gap),
    }
    if strengthened:
        e_long, _, ev_long, _ = long_optimize(factory, inst, [1.0] * p, [1.0] * p)
        g0, b0 = dep["angles"]
        e_warm, _, ev_warm, _ = long_optimize(factory, inst, list(g0), list(b0))
        out.update({
            "long_E": e_long,
            "long_evals": ev_long,
            "warm_E": min(e_warm, dep["E"]),
            "warm_evals": dep["evals"] + ev_warm,
        })
    return out


def run_experiment(train_rows=1_000_000, train_instances=2048, sizes=(16, 18, 20), p=4, epochs=20, test_n=20,
                   test_instances=8, seed=2024, top=32, progress=None):
    """The learned-proposal experiment end to end: generate the data set with the batched engine,
    train the surrogate, compare learned proposals with the plain optimiser
    on new instances (compare_on_instance). Returns a JSON-ready summary."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    X, Y, t_eval = generate_dataset(list(sizes), train_instances, train_rows, p, rng, progress=progress)
    t_gen = time.perf_counter() - t0
    model = Surrogate(X.shape[1])
    t0 = time.perf_counter()
    losses = model.train(X, Y, epochs=epochs)
    t_train = time.perf_counter() - t0
    if progress:
        progress(f"trained: {epochs} epochs in {t_train:.1f} s, loss {losses[0]:.4g} -> {losses[-1]:.4g}")
    results = []
    for t in range(test_instances):
        inst = random_ising(test_n, rng)
        results.append(compare_on_instance(model, inst, p, rng, top=top))
        if progress:
            progress(f"test instance {t}: {json.dumps(results[-1])}")
    model.close()
    met = [r for r in results if r["within_gap"] and r["eval_ratio"] >= 5]
    return {"rows": int(X.shape[0]), "gen_s": round(t_gen, 1), "gpu_eval_s": round(t_eval, 1),
            "train_s": round(t_train, 1), "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
            "within_gap": f"{sum(r['within_gap'] for r in results)}/{len(results)}",
            "target_met": f"{len(met)}/{len(results)}",
            "median_eval_ratio": float(np.median([r["eval_ratio"] for r in results])),
            "results": results}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-rows", type=int, default=1_000_000)
    ap.add_argument("--train-instances", type=int, default=2048)
    ap.add_argument("--sizes", default="16,18,20")
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--test-n", type=int, default=20)
    ap.add_argument("--test-instances", type=int, default=8)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--top", type=int, default=32)
    args = ap.parse_args()

    def progress(text):
        print(text, file=sys.stderr, flush=True)

    out = run_experiment(args.train_rows, args.train_instances, [int(x) for x in args.sizes.split(",")], args.p,
                         args.epochs, args.test_n, args.test_instances, args.seed, args.top, progress)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
