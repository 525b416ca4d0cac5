"""Readout sampling restated as plain Python loops -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of mossquantum_amd.readout.

Each function takes its uniforms as an explicit list, so a test can hand the
same uniforms to the product and to this restatement.

  draw_samples     reference _draw_samples!  src/execution/probabilities.jl:107-121
  discrete_sample  reference discrete_sample src/execution/sampling.jl:3-18

Pins (tests/test_readout.py): a point mass draws its index for every
uniform; hand-worked draws on p = (0.1, 0.2, 0.3, 0.4), including a uniform
that lands exactly on a cumulative boundary (searchsortedlast counts entries
<= x, so the boundary belongs to the next bin); trailing zero weights are
never drawn; every draw lies in its cumulative interval (brute force over
random distributions).
"""
from __future__ import annotations


def cumsum(p):
    """Left-to-right running sum (Julia cumsum on a Vector{Float64})."""
    out, acc = [], 0.0
    for x in p:
        acc += float(x)
        out.append(acc)
    return out


def searchsortedlast(c, x):
    """1-based index of the last entry <= x in sorted c, 0 if none."""
    last = 0
    for i, v in enumerate(c, start=1):
        if v <= x:
            last = i
    return last


def draw_samples(uniforms, p, nbits):
    """List of shots, each a list of nbits readout bits."""
    p = [float(x) for x in p]
    if not all(x == x and x not in (float("inf"),) and x >= 0 for x in p):
        raise ValueError("invalid sampling distribution")
    c = cumsum(p)
    total = c[-1]
    if not (total == total and total != float("inf") and total > 0):
        raise ValueError("sampling probabilities must have positive finite sum")
    last_positive = max(i for i, x in enumerate(p, start=1) if x > 0)
    shots = []
    for u in uniforms:
        selected = searchsortedlast(c, u * total) + 1
        basis = (last_positive if selected > len(p) else selected) - 1
        shots.append([(basis >> bit) & 1 for bit in range(nbits)])
    return shots


def discrete_sample(u, w):
    """1-based index for one uniform u."""
    w = [float(x) for x in w]
    total = 0.0
    for x in w:
        total += x
    threshold = u * total
    acc, last_positive = 0.0, 1
    for i, x in enumerate(w, start=1):
        if x > 0:
            last_positive = i
        acc += x
        if threshold < acc:
            return i
    return last_positive
