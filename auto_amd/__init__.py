"""auto_amd: MI355X (gfx950) Lenia world-step engine.

The engine is the C-ABI library libauto_hip.so (include/auto_hip.h) built
from auto_amd/csrc/ with hipcc.  This package mirrors the reference's
rule/step interface on top of it:

  automata.build(id, w, h)  ->  rule descriptor   (auto-core automata::build)
  Sim(desc, seed)           ->  engine handle     (auto-core Sim::new)
  sim.step(n), sim.wait(), sim.read_field(name), sim.write_field(name, data)

There is no CPU fallback: creating a Sim without the built library raises.
"""
from . import workloads
from .sim import Sim, Batch, Surrogate, build, available

__all__ = ["workloads", "Sim", "Batch", "Surrogate", "build", "available"]
