"""Gradient-memory strategies: ``Full()``, ``BoundarySaving()``, ``Ckpt()``.

Mirrors the reference's ``propagator/options.py:186-411,555-600``: the same
names, fields and validation, for the parts this build implements.

* ``Full()`` keeps the per-step imaging field ``v^2 w_sum`` of every step.
* ``BoundarySaving(storage='gpu')`` keeps the ``M+1``-wide strips around the
  physical box per step plus the last two wavefields, and reconstructs the
  forward field backwards in time (the ``impl='c'`` default).
* ``Ckpt(mode='chunk', chunks=c)`` keeps ``c`` state snapshots and replays
  each chunk before reversing it.  ``mode='recursive'`` with ``count`` slots
  keeps ``count`` snapshots and nothing else, and reverses on a binomial
  schedule: with r the smallest integer with C(count + r, count) >= nt, every
  step is re-run at most r times (the reference bisects recursively).

``BoundarySaving(storage_dtype='bf16')`` keeps the strips in bfloat16 (round
to nearest even), half the bytes of fp32.  Host/disk staging, fp16/int8 or
compressed strip storage and tail truncation are storage tiers outside this
build; asking for them raises.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar


@dataclass
class Full:
    """Keep everything the backward needs in memory."""
    strategy: ClassVar[str] = "full"


@dataclass
class BoundarySaving:
    """Reconstruct the forward wavefield from saved boundary strips."""
    storage: str = "gpu"
    storage_dtype: str = "fp32"
    transfer_interval: int | None = None
    pinned_memory: bool | None = None
    disk_dir: str | None = None
    ring_buffers: int | None = None
    disk_async_read: bool = False
    tail_steps: int | None = None
    compress: str | None = None
    strategy: ClassVar[str] = "boundary"

    def __post_init__(self):
        if self.storage not in {"gpu", "cpu", "disk"}:
            raise ValueError("BoundarySaving.storage must be 'gpu', 'cpu' or 'disk'.")
        if self.storage_dtype not in {"fp32", "fp16", "bf16", "int8"}:
            raise ValueError("BoundarySaving.storage_dtype must be fp32, fp16, bf16 or int8.")

    def check_supported(self):
        if self.storage != "gpu":
            raise NotImplementedError(
                f"BoundarySaving(storage={self.storage!r}): host/disk staging is not part of this "
                "build; the strips stay in HBM (288 GB per MI355X).")
        if self.storage_dtype not in ("fp32", "bf16"):
            raise NotImplementedError(
                f"BoundarySaving(storage_dtype={self.storage_dtype!r}) is not built; use fp32 or bf16.")
        if self.tail_steps or self.compress:
            raise NotImplementedError("BoundarySaving tail_steps / compress are not built.")


@dataclass
class Ckpt:
    """Rematerialise the forward wavefield from checkpoints."""
    mode: str = "chunk"
    chunks: int = 100
    count: int = 0
    storage: str = "gpu"
    pinned_memory: bool | None = None
    strategy: ClassVar[str] = "ckpt"

    def __post_init__(self):
        if self.storage not in {"gpu", "cpu"}:
            raise ValueError("Ckpt.storage must be 'gpu' or 'cpu'.")
        if self.mode == "chunk":
            if self.chunks < 1:
                raise ValueError("Ckpt.chunks must be >= 1 when mode='chunk'.")
            if self.count != 0:
                raise ValueError("Ckpt.count is only valid when mode='recursive'.")
        elif self.mode == "recursive":
            if self.count < 1:
                raise ValueError("Ckpt.count must be >= 1 when mode='recursive'.")
            if self.chunks != 100:
                raise ValueError("Ckpt.chunks is only valid when mode='chunk'.")
        else:
            raise ValueError("Ckpt.mode must be 'chunk' or 'recursive'.")

    def interval(self, nt: int) -> int:
        """Steps between snapshots for an ``nt``-step run."""
        slots = self.chunks if self.mode == "chunk" else self.count
        slots = max(1, min(int(slots), max(nt, 1)))
        return max(1, -(-nt // slots))


def as_memory_strategy(value):
    if value is None or isinstance(value, (Full, BoundarySaving, Ckpt)):
        return value
    if isinstance(value, dict):
        kind = value.get("kind", value.get("strategy"))
        rest = {k: v for k, v in value.items() if k not in ("kind", "strategy")}
        if kind == "full":
            return Full()
        if kind == "boundary":
            return BoundarySaving(**rest.get("boundary", rest))
        if kind == "ckpt":
            return Ckpt(**rest.get("ckpt", rest))
    raise TypeError(f"not a memory strategy: {value!r}")
