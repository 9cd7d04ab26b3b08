"""Random states that a run can reach, for the parity tests.

Fields drawn independently at random are not a state any run reaches: after
every step of the reference, E = M . D holds in each cell whose E the D/E
update produces (fdtd.c:3404-3462).  The FP32 precision form relies on it
and carries E alone for those cells (E^{n+1} = E^n + M . dD).  reachable()
advances a random state by one step of the oracle, after which the relation
holds, and returns every field and slot state for the test to write into the
simulation and into the oracle it compares with.
"""
from oracle import orc


def reachable(spec, medium, fields, states):
    """(fields, states) one oracle step after the given ones: fields is the
    list of 9 components in file order, states maps (material, kind) to the
    slot-state vectors of phpfdtd_write_state."""
    o = orc.Oracle(spec, medium)
    for comp, f in enumerate(fields):
        o.set_field(comp, f)
    for (m, kind), a in states.items():
        o.set_state(m, kind, a)
    o.step(1)
    out_f = [o.field(c) for c in range(9)]
    out_s = {k: o.state(*k) for k in states}
    o.close()
    return out_f, out_s
