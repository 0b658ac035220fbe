"""GEMM autotuner of the library: the distinct MFMA complex-GEMM shapes of a workload are
recorded while it runs once, every kernel variant (K tile 16/32 x LDS row padding 1/16)
and split-K factor is timed on each shape (dc2_gemm_tune), and the fastest choice that
beats the built-in rule is installed in the library's process-wide table, which the GEMM
dispatch consults for products of exactly that shape.  A table can be saved as JSON and
installed again in another process on the same hardware (load)."""
import json
from typing import Callable, Dict, List

from .api import Caspt2Context, gemm_clear_choices, gemm_set_choice


def shapes_of(ctx: Caspt2Context, run: Callable[[], None]) -> List[tuple]:
    """The distinct MFMA GEMM shapes (opA, opB, M, N, K, batch) that run() launches on ctx."""
    ctx.set_profiling(True)
    try:
        run()
        ctx.synchronize()
        return ctx.gemm_shapes()
    finally:
        ctx.set_profiling(False)


def tune(ctx: Caspt2Context, shapes: List[tuple], reps: int = 3) -> List[Dict]:
    """Tune every shape; returns one record per shape (the choice and both timings)."""
    table = []
    for (oa, ob, M, N, K, batch) in shapes:
        r = ctx.gemm_tune(oa, ob, M, N, K, batch, reps)
        table.append({"opA": oa, "opB": ob, "M": M, "N": N, "K": K, "batch": batch, **r})
    return table


def tune_workload(ctx: Caspt2Context, run: Callable[[], None], reps: int = 3) -> Dict:
    """Record run()'s shapes, tune them, and summarise: shapes, how many got a new choice,
    and the summed per-launch time of the built-in choices against the tuned ones."""
    table = tune(ctx, shapes_of(ctx, run), reps)
    changed = [t for t in table if t["variant"] > 0]
    return {"table": table, "shapes": len(table), "changed": len(changed),
            "builtin_us": 1e6 * sum(t["builtin_seconds"] for t in table),
            "tuned_us": 1e6 * sum(t["seconds"] for t in table)}


def save(table: List[Dict], path: str) -> None:
    with open(path, "w") as f:
        json.dump([t for t in table if t["variant"] > 0], f, indent=1)


def load(path: str, clear: bool = True) -> int:
    """Install the choices of a saved table; returns how many."""
    if clear:
        gemm_clear_choices()
    with open(path) as f:
        table = json.load(f)
    for t in table:
        gemm_set_choice(t["opA"], t["opB"], t["M"], t["N"], t["K"], t["batch"], t["variant"], t["split"])
    return len(table)
