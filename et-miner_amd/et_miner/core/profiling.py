"""Per-phase wall times of a mining call (the reference's ``core/profiling.py:26-177``, without RSS sampling)."""

from __future__ import annotations

import time
from dataclasses import dataclass, field


@dataclass
class Phase:
    name: str
    duration_ms: float
    extra: dict = field(default_factory=dict)


class ProfilingSession:
    def __init__(self) -> None:
        self.phases: list = []
        self._open: tuple | None = None

    def start_phase(self, name: str) -> None:
        self._open = (name, time.perf_counter())

    def end_phase(self, **extra) -> None:
        if self._open is None:
            return
        name, t0 = self._open
        self.phases.append(Phase(name, (time.perf_counter() - t0) * 1000, dict(extra)))
        self._open = None

    def record_phase(self, name: str, duration_ms: float, **extra) -> None:
        self.phases.append(Phase(name, float(duration_ms), dict(extra)))

    def total_ms(self) -> float:
        return sum(p.duration_ms for p in self.phases)

    def summary(self) -> str:
        return "\n".join(f"{p.name:>24s} {p.duration_ms:10.2f} ms {p.extra}" for p in self.phases)
