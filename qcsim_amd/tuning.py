"""Autotuner cache: launch variants chosen by qcs_autotune, saved per device.

qcs_autotune() times every launch variant of the streaming gate kernels for
each target qubit and keeps the fastest for the process. This module saves
those choices as JSON keyed by (architecture, precision, qubit count) and
presets them in later processes, so benchmarks and profiles run the same
kernels without re-tuning. The cache shipped in tune_cache/ was measured on
an MI355X.

    python -m qcsim_amd.tuning --qubits 30 --precision c128 --out FILE
"""
from __future__ import annotations

import argparse
import json
import os

from ._lib import QCS_COMPLEX64, QCS_COMPLEX128, check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(HERE, "tune_cache", "gfx950.json")
_PREC = {"c128": QCS_COMPLEX128, "c64": QCS_COMPLEX64}


def _key(precision: str, n: int) -> str:
    return f"{precision}:{n}"


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def apply_cache(precision: str, n: int, path: str = DEFAULT_CACHE) -> bool:
    """Preset the engine's variant table for (precision, n); True if found."""
    entry = load_cache(path).get(_key(precision, n))
    if not entry:
        return False
    for q, vid in enumerate(entry["variant"]):
        check(lib().qcs_set_tuned_variant(_PREC[precision], n, q, int(vid)))
    return True


def tune(reg, reps: int = 3) -> dict:
    """Run the autotuner on `reg` and return its cache entry."""
    gbps = reg.autotune(reps)
    prec = _PREC[reg.precision]
    variant = [int(lib().qcs_get_tuned_variant(prec, reg.NrQubits, q)) for q in range(reg.NrQubits)]
    return {"variant": variant, "best_gbps": [round(float(x), 1) for x in gbps.max(axis=1)]}


def ensure_tuned(reg, path: str = DEFAULT_CACHE) -> str:
    """Preset from the cache when it has this (precision, n); else tune now."""
    if apply_cache(reg.precision, reg.NrQubits, path):
        return "cache"
    tune(reg)
    return "measured"


def main():
    from .register import QubitRegister

    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--precision", default="c128", choices=sorted(_PREC))
    ap.add_argument("--reps", type=int, default=4)
    ap.add_argument("--out", default=DEFAULT_CACHE)
    args = ap.parse_args()
    reg = QubitRegister(args.qubits, precision=args.precision)
    reg.setToEqualSuperposition()
    entry = tune(reg, args.reps)
    cache = load_cache(args.out)
    cache[_key(args.precision, args.qubits)] = entry
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(cache, f, indent=1)
    print(json.dumps({_key(args.precision, args.qubits): entry}))


if __name__ == "__main__":
    main()
