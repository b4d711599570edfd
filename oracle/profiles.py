"""CPU oracle of the mean profiles (MeanProfiles, logging.jl:310-525); test
infrastructure only.

Samples (log_sample!/hmean, logging.jl:380-413): the real part of the (0,0)
mode of vel1, vel2 (C-nodes) and vel3 (interior I-nodes, 0 in the two
boundary entries); plane means of the advection terms adv1-3 (adv3 with 0
boundary entries) and of the Smagorinsky stresses sgs11, sgs12, sgs22, sgs33
(C-nodes) and sgs13, sgs23 (I-nodes, the wall stresses as boundary entries),
all at the rate evaluation of the current state.  The time average is the
trapezoidal rule over the samples after each step (prepare_samples!,
process_log!, logging.jl:370-451).
"""

import numpy as np

from .model import compute_rates

NAMES = ("vel1", "vel2", "vel3", "adv1", "adv2", "adv3", "sgs11", "sgs12", "sgs13", "sgs22", "sgs23", "sgs33")


def samples(model, state, names=NAMES):
    terms = {}
    rates = {f: np.zeros_like(a) for f, a in state.items()}
    compute_rates(model, rates, state, terms)
    out = {}
    pad = lambda a: np.concatenate([[0.0], a, [0.0]])
    for n in names:
        if n in ("vel1", "vel2"):
            out[n] = state[n][0, 0, :].real.copy()
        elif n == "vel3":
            out[n] = pad(state["vel3"][0, 0, :].real)
        elif n == "adv3":
            out[n] = pad(terms[n].mean(axis=(0, 1)))
        else:
            out[n] = terms[n].mean(axis=(0, 1))
    return out


def trapezoid(sample_list, dt):
    """∫ s dt over equally spaced samples (the first one at the start time)."""
    acc = {n: np.zeros_like(v) for n, v in sample_list[0].items()}
    for a, b in zip(sample_list[:-1], sample_list[1:]):
        for n in acc:
            acc[n] += dt / 2 * (a[n] + b[n])
    return acc
